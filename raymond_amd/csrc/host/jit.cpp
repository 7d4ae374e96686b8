// Run-time compilation of a scene's material code with hiprtc.
//
// Replaces the reference's per-scene Metal JIT (MaterialBuilder.swift:253-325)
// and the include relocation step (buildtools/jit-relocate.rb): the device
// headers are embedded in this library and handed to hiprtc as in-memory
// includes.  Code objects are cached on disk, keyed by a hash of the full
// translation unit, the options and the header texts.
#include <dlfcn.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "core.h"
#include "embedded_headers.inc"

namespace rm {

std::string jit_translation_unit(const std::string& body) {
    std::string tu;
    tu += "#include \"material_prelude.h\"\n";
    tu += body;
    tu += "\n#include \"shade_kernels.h\"\n";
    return tu;
}

static uint64_t fnv1a(const void* data, size_t n, uint64_t h) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

// `jitcache/` beside libraymond_amd.so: code objects compiled ahead of time
// (rmh_jit_precompile with RM_JIT_CACHE pointing there) travel with the build.
static std::string bundled_cache_dir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&bundled_cache_dir), &info) && info.dli_fname) {
        std::string lib = info.dli_fname;
        const size_t slash = lib.rfind('/');
        if (slash != std::string::npos) return lib.substr(0, slash) + "/jitcache";
    }
    return "jitcache";
}

static std::string cache_dir() {
    if (const char* e = getenv("RM_JIT_CACHE")) return e;
    if (const char* home = getenv("HOME")) {
        std::string d = std::string(home) + "/.cache";
        mkdir(d.c_str(), 0755);
        d += "/raymond_amd";
        mkdir(d.c_str(), 0755);
        if (access(d.c_str(), W_OK) == 0) return d;
    }
    std::string d = "/tmp/raymond_amd_jit_" + std::to_string(getuid());
    mkdir(d.c_str(), 0755);
    return d;
}

static std::vector<std::string> split_options(const std::string& s) {
    std::vector<std::string> out;
    std::istringstream is(s);
    std::string tok;
    while (is >> tok) out.push_back(tok);
    return out;
}

bool jit_compile(const std::string& body, const std::string& options, std::vector<char>& code, std::string& log,
                 bool& cache_hit) {
    cache_hit = false;
    const std::string tu = jit_translation_unit(body);
    std::vector<std::string> opts = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                                     "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang",
                                     "-packed-fp32-ops"};
    for (auto& o : split_options(options)) opts.push_back(o);
    // extra compiler options for experiments (part of the cache key like the rest)
    if (const char* e = getenv("RM_JIT_OPTIONS"))
        for (auto& o : split_options(e)) opts.push_back(o);

    uint64_t h = 1469598103934665603ull;
    h = fnv1a(tu.data(), tu.size(), h);
    for (auto& o : opts) h = fnv1a(o.data(), o.size() + 1, h);
    for (int i = 0; i < kEmbeddedHeaderCount; i++) h = fnv1a(kEmbeddedHeaderTexts[i], strlen(kEmbeddedHeaderTexts[i]), h);
    char name[64];
    snprintf(name, sizeof(name), "%016llx.co", (unsigned long long)h);
    const std::string path = cache_dir() + "/" + name;
    // the writable cache first, then a read-only one shipped next to the library
    for (const std::string& p : {path, bundled_cache_dir() + "/" + name}) {
        std::ifstream f(p, std::ios::binary);
        if (f) {
            code.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
            if (!code.empty()) {
                cache_hit = true;
                return true;
            }
        }
    }

    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, tu.c_str(), "materials.hip", kEmbeddedHeaderCount, kEmbeddedHeaderTexts,
                            kEmbeddedHeaderNames) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    std::vector<const char*> copts;
    for (auto& o : opts) copts.push_back(o.c_str());
    const hiprtcResult r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    if (log_size > 1) {
        log.resize(log_size);
        hiprtcGetProgramLog(prog, &log[0]);
        while (!log.empty() && log.back() == '\0') log.pop_back();
    }
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        if (log.empty()) log = hiprtcGetErrorString(r);
        return false;
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    code.resize(code_size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);

    const std::string tmp = path + ".tmp" + std::to_string(getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        f.write(code.data(), (std::streamsize)code.size());
    }
    rename(tmp.c_str(), path.c_str());
    return true;
}

}  // namespace rm

extern "C" int rmh_jit_precompile(const char* source, size_t length, const char* options, int* out_hit, char* log,
                                  size_t log_capacity) {
    if (out_hit) *out_hit = 0;
    if (!source) return RM_ERROR_INVALID;
    std::vector<char> code;
    std::string text;
    bool hit = false;
    const bool ok = rm::jit_compile(std::string(source, length), options ? options : "", code, text, hit);
    if (log && log_capacity) {
        const size_t n = text.size() < log_capacity - 1 ? text.size() : log_capacity - 1;
        memcpy(log, text.data(), n);
        log[n] = '\0';
    }
    if (out_hit) *out_hit = hit ? 1 : 0;
    return ok ? RM_OK : RM_ERROR_COMPILE;
}
