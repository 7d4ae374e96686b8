"""Builds raymond_amd/libraymond_amd.so (HIP kernels for gfx950 + host runtime).

hipcc compiles the .hip kernel sources; g++ compiles the host runtime against
the HIP headers; hipcc links the shared library with hiprtc.  Objects go to
raymond_amd/_build/.  Every build recompiles (the driver's build check).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libraymond_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
]

KERNELS = ["kernels/trace.hip", "kernels/misc.hip", "kernels/sort.hip", "kernels/rcache.hip", "kernels/noalloc.hip"]
HOST = ["host/bvh_build.cpp", "host/jit.cpp", "host/material_builder.cpp", "host/bvh_check.cpp", "host/scene.cpp", "host/renderer.cpp", "host/rcache.cpp", "host/batch.cpp",
        "host/capi.cpp", "host/ply_reader.cpp", "host/exr_writer.cpp", "host/sky.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(line for line in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in line and "not a recognized feature" not in line)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


# Test build of the library with one LDS row of traversal stack, so that the
# spill rows carry every deeper stack entry (tests/test_full_size_gpu.py).
# The sources that include kernels/trace_api.h are compiled again with the
# define; the other objects are shared with the main build.
SPILL_LIB = os.path.join(HERE, "libraymond_amd_spill.so")
SPILL_DEFINES = ["-DRM_LDS_STACK=1"]
SPILL_SOURCES = ["kernels/trace.hip", "kernels/noalloc.hip", "host/scene.cpp", "host/renderer.cpp", "host/batch.cpp",
                 "host/capi.cpp"]


def _compile(src, defines=(), tag=""):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, tag + src.replace("/", "_") + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + list(defines) + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + HOST_FLAGS + list(defines) + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


def _link(objs, out):
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lhiprtc", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, out)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    _run([sys.executable, os.path.join(CSRC, "embed_headers.py"), os.path.join(CSRC, "device"),
          os.path.join(CSRC, "host", "embedded_headers.inc")])
    sources = KERNELS + HOST
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, sources))
        spill_objs = list(ex.map(lambda src: _compile(src, SPILL_DEFINES, "spill_"), SPILL_SOURCES))
    _link(objs, LIB)
    by_src = dict(zip(sources, objs))
    by_src.update(zip(SPILL_SOURCES, spill_objs))
    _link([by_src[src] for src in sources], SPILL_LIB)
    if verbose:
        print("built", LIB, "and", SPILL_LIB)
    return LIB


JIT_CACHE = os.path.join(HERE, "jitcache")


def warm_jit_cache(verbose=False, big=False):
    """Compiles the material modules of the scenes the tests, smoke check and
    benchmark workloads load into jitcache/ beside the library, which the JIT
    reads before compiling; `big` adds the 32-graph scene's out-of-line form.
    hiprtc serialises compilations inside one process, so each module gets a
    process of its own.  Returns {scene: seconds}."""
    import json
    root = os.path.dirname(HERE)
    sys.path.insert(0, root)
    from raymond_amd import scenes
    jobs = scenes.bundled_material_sources(big=big)
    os.makedirs(JIT_CACHE, exist_ok=True)
    env = dict(os.environ, RM_JIT_CACHE=JIT_CACHE)
    code = ("import sys, time; sys.path.insert(0, %r); import raymond_amd.api as api; "
            "src = sys.stdin.read(); t = time.time(); api.jit_precompile(src, sys.argv[1]); "
            "print(round(time.time() - t, 1))" % root)
    # batches compile their library again with the batch shading kernel
    by_name = dict(jobs)
    jobs = [(name, src, "") for name, src in jobs] + [
        ("cornell, batch kernel", by_name["cornell"], "-DRM_BATCH_KERNEL"),
        ("lamp box, batch kernel", by_name["lamp box"], "-DRM_BATCH_KERNEL"),
        ("cornell, cache kernel", by_name["cornell"], "-DRM_CACHE_KERNEL")]
    procs = []
    for name, src, options in jobs:
        p = subprocess.Popen([sys.executable, "-c", code, options], env=env, stdin=subprocess.PIPE,
                             stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        procs.append((name, src, p))
    times = {}
    for name, src, p in procs:
        out, err = p.communicate(src)
        if p.returncode != 0:
            raise RuntimeError("precompiling %s failed:\n%s" % (name, err[-4000:]))
        times[name] = float(out.strip().splitlines()[-1])
    if verbose:
        print("jit cache:", json.dumps(times))
    return times


if __name__ == "__main__":
    build(verbose=True)
    if "--warm" in sys.argv:
        warm_jit_cache(verbose=True, big="--big" in sys.argv)
