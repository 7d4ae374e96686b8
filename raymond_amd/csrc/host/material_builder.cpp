// Material code generation: Cycles-style node graphs (the Blender exporter's
// JSON, scripts/blender_exporter/raymond_blender/materials.py:300-560) to
// source text defining shade_surface / shade_light.
//
// Restates host/scene/MaterialBuilder.swift:32-826: depth-first emission from
// the output nodes (emitNode :793-825), per-node invocations with template
// parameters (:327-662), default inputs (:762-791), camelCase identifiers
// (:19-43) and the per-table switch functions (:263-300).  The same text is
// compiled by hiprtc for the GPU and by g++ for the CPU oracle.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <algorithm>
#include <cctype>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/rm_host.h"
#include "json.h"

namespace rm {
namespace sky {
void precompute_sun(float sun_elevation, float angular_diameter, float altitude, float air_density,
                    float dust_density, float* bottom, float* top);
}
namespace {

using json::Value;

struct CodegenError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// "Base Color" -> "baseColor": words capitalised, spaces removed, first letter
// lower-cased, non-identifier characters replaced by '_' (MaterialBuilder.swift:19-43).
std::string make_identifier(const std::string& name) {
    std::string trimmed = name;
    while (!trimmed.empty() && trimmed.front() == ' ') trimmed.erase(trimmed.begin());
    while (!trimmed.empty() && trimmed.back() == ' ') trimmed.pop_back();
    if (trimmed.empty()) return "";
    std::string cap;
    bool word_start = true;
    for (char c : trimmed) {
        if (c == ' ') {
            word_start = true;
            continue;
        }
        if (word_start && std::isalpha((unsigned char)c)) cap += (char)std::toupper((unsigned char)c);
        else cap += (char)std::tolower((unsigned char)c);
        word_start = false;  // words are delimited by spaces only (NSString.capitalized)
    }
    std::string out;
    out += (char)std::tolower((unsigned char)trimmed[0]);
    out += cap.substr(1);
    for (char& c : out)
        if (!(std::isalnum((unsigned char)c) || c == '_')) c = '_';
    return out;
}

std::string enum_token(const std::string& v) {
    std::string out;
    for (char c : v) out += std::isalnum((unsigned char)c) || c == '_' ? (char)std::toupper((unsigned char)c) : '_';
    return out;
}

std::string fmt_float(double v) {
    if (std::isnan(v)) return "__builtin_nanf(\"\")";
    if (std::isinf(v)) return v > 0 ? "__builtin_huge_valf()" : "(-__builtin_huge_valf())";
    char buf[64];
    snprintf(buf, sizeof(buf), "%.9g", (double)(float)v);
    std::string s = buf;
    if (s.find_first_of(".eE") == std::string::npos) s += ".0";
    return s + "f";
}

std::string literal(const Value& v) {
    if (v.is_number()) return fmt_float(v.num);
    if (v.kind == Value::Bool) return v.b ? "1.0f" : "0.0f";
    if (v.is_array()) {
        const size_t n = v.arr.size();
        if (n < 1 || n > 4) throw CodegenError("unsupported vector size");
        if (n == 1) return fmt_float(v.arr[0].num);
        std::string s = "vec" + std::to_string(n) + "(";
        for (size_t i = 0; i < n; i++) s += (i ? ", " : "") + fmt_float(v.arr[i].num);
        return s + ")";
    }
    throw CodegenError("unsupported input value");
}

std::string coercion(const std::string& socket_type) {
    if (socket_type == "RGBA") return "RGBA";
    if (socket_type == "VECTOR") return "VECTOR";
    if (socket_type == "SHADER") return "SHADER";
    return "VALUE";  // VALUE, INT, BOOLEAN
}

const Value* param(const Value& node, const char* key) {
    const Value* p = node.get("parameters");
    return p ? p->get(key) : nullptr;
}
std::string param_str(const Value& node, const char* key, const char* def) {
    const Value* v = param(node, key);
    return v && v->is_string() ? v->str : def;
}
double param_num(const Value& node, const char* key, double def) {
    const Value* v = param(node, key);
    if (v && v->is_number()) return v->num;
    if (v && v->kind == Value::Bool) return v->b ? 1.0 : 0.0;
    return def;
}
bool param_bool(const Value& node, const char* key) {
    const Value* v = param(node, key);
    return v && ((v->kind == Value::Bool && v->b) || (v->is_number() && v->num != 0));
}

struct Invocation {
    std::string kernel;                      // full C++ type
    std::map<std::string, std::string> in;   // field -> expression (sorted like the reference)
    std::vector<std::string> extra;          // raw statements after the assignments
    std::string comment;
};

// A texture slot of the material library: an image file, or the Nishita
// sky texture of the given atmosphere (512 x 512, MaterialBuilder.swift:351-372).
struct TextureSlot {
    std::string kind;  // "image" or "nishita"
    std::string path;
    float sun_elevation = 0.0f, altitude = 0.0f, air_density = 0.0f, dust_density = 0.0f, ozone_density = 0.0f;
    bool same(const TextureSlot& o) const {
        return kind == o.kind && path == o.path && sun_elevation == o.sun_elevation && altitude == o.altitude &&
               air_density == o.air_density && dust_density == o.dust_density && ozone_density == o.ozone_density;
    }
};

std::string json_string(const std::string& s) {
    std::string out = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') out += '\\';
        if ((unsigned char)c < 0x20) {
            char buf[8];
            snprintf(buf, sizeof(buf), "\\u%04x", (unsigned)(unsigned char)c);
            out += buf;
            continue;
        }
        out += c;
    }
    return out + "\"";
}

class Codegen {
  public:
    const std::vector<TextureSlot>& textures() const { return textures_; }

    void add_table(const std::string& fn, const Value& materials) {
        if (!materials.is_array()) throw CodegenError("material table must be an array");
        std::vector<std::string> fns;
        for (const Value& m : materials.arr) {
            const Value* name = m.get("name");
            const Value* nodes = m.get("nodes");
            if (!name || !name->is_string() || !nodes || !nodes->is_object())
                throw CodegenError("material needs 'name' and 'nodes'");
            fns.push_back(add_material(*nodes, name->str));
        }
        tables_.push_back({fn, fns});
    }

    std::string build(bool noinline) const {
        std::string out;
        out += "// generated material code\n";
        out += "namespace rm {\n\n";
        const std::string q = noinline ? "RM_MATERIAL_FN_NOINLINE " : "RM_HD ";
        for (const auto& f : functions_) out += q + "void " + f.first + "(const SceneView& ctx, ShadingContext& shading);\n";
        out += "\n";
        for (const auto& t : tables_) {
            out += "RM_HD void " + t.first + "(uint32_t index, const SceneView& ctx, ShadingContext& shading) {\n";
            out += "    switch (index) {\n";
            for (size_t i = 0; i < t.second.size(); i++)
                out += "    case " + std::to_string(i) + ": " + t.second[i] + "(ctx, shading); break;\n";
            out += "    default: break;\n    }\n}\n\n";
        }
        for (const auto& f : functions_) out += (noinline ? "RM_MATERIAL_FN_NOINLINE" : "RM_HD") + f.second.substr(5);
        out += "}  // namespace rm\n";
        return out;
    }

  private:
    std::vector<TextureSlot> textures_;  // slots in first-use order
    std::vector<std::pair<std::string, std::string>> functions_;  // name, definition
    std::vector<std::pair<std::string, std::vector<std::string>>> tables_;
    std::set<std::string> used_names_;
    std::map<std::string, int> state_;  // 1 pending, 2 emitted
    std::vector<std::pair<std::string, Invocation>> invocations_;

    std::string unused_function_name(const std::string& name) {
        const std::string base = "s_" + make_identifier(name);
        std::string cand = base;
        for (int i = 2; used_names_.count(cand); i++) cand = base + std::to_string(i);
        used_names_.insert(cand);
        return cand;
    }

    static bool is_output(const std::string& t) {
        return t == "OUTPUT_MATERIAL" || t == "OUTPUT_WORLD" || t == "OUTPUT_LIGHT";
    }

    std::string add_material(const Value& nodes, const std::string& name) {
        state_.clear();
        invocations_.clear();
        std::vector<std::string> keys;
        for (const auto& kv : nodes.obj) keys.push_back(kv.first);
        std::sort(keys.begin(), keys.end());
        for (const auto& k : keys) {
            const Value* t = nodes.get(k)->get("type");
            if (t && t->is_string() && is_output(t->str)) emit_node(nodes, k);
        }
        const std::string fn = unused_function_name(name);
        std::string text = "RM_HD void " + fn + "(const SceneView& ctx, ShadingContext& shading) {\n";
        for (const auto& inv : invocations_) {
            const std::string var = "n_" + make_identifier(inv.first);
            text += "\n";
            if (!inv.second.comment.empty()) text += "    // " + inv.second.comment + "\n";
            text += "    " + inv.second.kernel + " " + var + "{};\n";
            for (const auto& kv : inv.second.in) text += "    " + var + "." + kv.first + " = " + kv.second + ";\n";
            for (const auto& e : inv.second.extra) text += "    " + var + "." + e + ";\n";
            text += "    " + var + ".compute(ctx, shading);\n";
        }
        text += "}\n\n";
        functions_.push_back({fn, text});
        return fn;
    }

    // Node types with a kernel of fixed form (MaterialBuilder.swift:442-603);
    // most are the reference's constant stubs.
    static const char* stub_kernel(const std::string& t) {
        static const std::map<std::string, const char*> kernels = {
            {"TEX_BRICK", "TexBrick"}, {"TEX_MAGIC", "TexMagic"}, {"TEX_MUSGRAVE", "TexMusgrave"},
            {"TEX_IES", "TexIES"}, {"AMBIENT_OCCLUSION", "AmbientOcclusion"}, {"VOLUME_SCATTER", "VolumeScatter"},
            {"VOLUME_ABSORPTION", "VolumeAbsorption"}, {"PRINCIPLED_VOLUME", "PrincipledVolume"},
            {"DISPLACEMENT", "Displacement"}, {"ATTRIBUTE", "Attribute"}, {"OBJECT_INFO", "ObjectInfo"},
            {"PARTICLE_INFO", "ParticleInfo"}, {"LIGHT_FALLOFF", "LightFalloff"}, {"VERTEX_COLOR", "VertexColor"},
            {"NORMAL", "NormalProduct"}, {"CURVE_RGB", "ColorCurves"}, {"BLACKBODY", "Blackbody"},
            {"UVMAP", "UVMapCoordinate"}};
        auto it = kernels.find(t);
        return it == kernels.end() ? nullptr : it->second;
    }

    int texture_slot(const TextureSlot& t) {
        for (size_t i = 0; i < textures_.size(); i++)
            if (textures_[i].same(t)) return (int)i;
        textures_.push_back(t);
        return (int)textures_.size() - 1;
    }

    // Image and environment textures (MaterialBuilder.swift:379-435): one slot
    // per file, linear interpolation, repeat extension, straight alpha
    Invocation image_texture(const Value& node, bool environment) {
        TextureSlot ts;
        ts.kind = "image";
        ts.path = param_str(node, "filepath", "");
        if (ts.path.empty()) throw CodegenError("image texture without a file");
        const std::string interp = enum_token(param_str(node, "interpolation", "Linear"));
        if (interp != "LINEAR") throw CodegenError("unsupported image interpolation '" + interp + "'");
        const std::string proj = enum_token(param_str(node, "projection", environment ? "EQUIRECTANGULAR" : "FLAT"));
        static const std::set<std::string> projections = {"FLAT", "BOX", "EQUIRECTANGULAR", "MIRROR_BALL"};
        if (!projections.count(proj)) throw CodegenError("unsupported image projection '" + proj + "'");
        const std::string cs = enum_token(param_str(node, "colorspace", "sRGB"));
        static const std::set<std::string> spaces = {"LINEAR", "SRGB", "NON_COLOR", "RAW", "XYZ", "FILMIC_LOG"};
        if (!spaces.count(cs)) throw CodegenError("unsupported colour space '" + cs + "'");
        Invocation inv;
        inv.kernel = "TexImage<" + std::to_string(texture_slot(ts)) + ", kTexImage::INTERPOLATION_" + interp +
                     ", kTexImage::PROJECTION_" + proj + ", kTexImage::EXTENSION_REPEAT, kTexImage::ALPHA_STRAIGHT, "
                     "kTexImage::COLOR_SPACE_" + cs + ">";
        inv.comment = ts.path;
        return inv;
    }

    // Sky texture (MaterialBuilder.swift:329-378): Nishita, or Hosek-Wilkie
    // mapped onto Nishita the reference's ad-hoc way
    Invocation sky_texture(const Value& node) {
        std::string type = enum_token(param_str(node, "sky_type", "NISHITA"));
        float elevation = (float)param_num(node, "sun_elevation", 0.2617994f);
        float rotation = (float)param_num(node, "sun_rotation", 0.0);
        bool disc = param(node, "sun_disc") ? param_bool(node, "sun_disc") : true;
        const float size = (float)param_num(node, "sun_size", 0.009512044f);
        const float intensity = (float)param_num(node, "sun_intensity", 1.0);
        const float altitude = (float)param_num(node, "altitude", 0.0);
        float air = (float)param_num(node, "air_density", 1.0);
        float dust = (float)param_num(node, "dust_density", 1.0);
        float ozone = (float)param_num(node, "ozone_density", 1.0);
        float scale = 1.0f;
        if (type == "HOSEK_WILKIE") {
            float d[3] = {0.0f, 0.0f, 1.0f};
            if (const Value* sd = param(node, "sun_direction"))
                if (sd->is_array() && sd->arr.size() == 3)
                    for (int k = 0; k < 3; k++) d[k] = (float)sd->arr[k].num;
            rotation = std::atan2(d[0], d[1]);
            elevation = std::max(std::asin(d[2]), 5.0f * 3.14159265358979323846f / 180.0f);
            disc = false;
            air = 2.0f;
            dust = 0.5f;
            ozone = 1.0f;
            scale = 0.08f;
            type = "NISHITA";
        }
        if (type != "NISHITA") throw CodegenError("unsupported sky type '" + type + "'");
        TextureSlot ts;
        ts.kind = "nishita";
        ts.sun_elevation = elevation;
        ts.altitude = altitude;
        ts.air_density = air;
        ts.dust_density = dust;
        ts.ozone_density = ozone;
        float data[10];
        sky::precompute_sun(elevation, size, altitude, air, dust, data, data + 3);
        data[6] = elevation;
        data[7] = rotation;
        data[8] = disc ? size : -1.0f;
        data[9] = intensity;
        Invocation inv;
        inv.kernel = "TexNishita<" + std::to_string(texture_slot(ts)) + ">";
        inv.in["scale"] = fmt_float(scale);
        for (int k = 0; k < 10; k++) inv.extra.push_back("data[" + std::to_string(k) + "] = " + fmt_float(data[k]));
        return inv;
    }

    // kernel type and fixed parameters per node type (MaterialBuilder.swift:327-662)
    Invocation make_invocation(const Value& node) {
        const Value* tv = node.get("type");
        if (!tv || !tv->is_string()) throw CodegenError("node without type");
        const std::string& t = tv->str;
        Invocation inv;
        if (t == "BSDF_PRINCIPLED") {
            inv.kernel = "BsdfPrincipled<kBsdfPrincipled::DISTRIBUTION_GGX, kBsdfPrincipled::SUBSURFACE_METHOD_" +
                         enum_token(param_str(node, "subsurface_method", "RANDOM_WALK")) + ">";
        } else if (t == "BSDF_DIFFUSE") {
            inv.kernel = "BsdfDiffuse";
        } else if (t == "EMISSION") {
            inv.kernel = "Emission";
        } else if (t == "BACKGROUND") {
            inv.kernel = "Background";
        } else if (t == "OUTPUT_MATERIAL") {
            inv.kernel = "OutputMaterial";
        } else if (t == "OUTPUT_WORLD") {
            inv.kernel = "OutputWorld";
        } else if (t == "OUTPUT_LIGHT") {
            inv.kernel = "OutputLight";
        } else if (t == "VALUE") {
            inv.kernel = "Value";
            const Value* v = param(node, "value");
            inv.in["value"] = fmt_float(v && v->is_number() ? v->num : 0.0);
        } else if (t == "RGB") {
            inv.kernel = "RGB";
            const Value* v = param(node, "color");
            if (!v) v = param(node, "value");
            inv.in["color"] = v ? literal(*v) : "vec4(0.0f)";
        } else if (t == "TEX_COORD") {
            inv.kernel = "TextureCoordinate";
        } else if (t == "NEW_GEOMETRY") {
            inv.kernel = "NewGeometry";
        } else if (t == "TEX_NOISE") {
            inv.kernel = "TexNoise<kTexNoise::DIMENSION_" + enum_token(param_str(node, "noise_dimensions", "3D")) + ">";
        } else if (t == "TEX_VORONOI") {
            inv.kernel = "TexVoronoi";
        } else if (t == "VALTORGB") {
            const Value* els = param(node, "elements");
            if (!els || !els->is_array() || els->arr.empty()) throw CodegenError("color ramp without elements");
            inv.kernel = "ColorRamp<" + std::to_string(els->arr.size()) + ">";
            for (size_t i = 0; i < els->arr.size(); i++) {
                const Value& e = els->arr[i];
                const Value* p = e.get("position");
                const Value* c = e.get("color");
                if (!p || !c) throw CodegenError("color ramp element needs position and color");
                inv.extra.push_back("elements[" + std::to_string(i) + "].position = " + literal(*p));
                inv.extra.push_back("elements[" + std::to_string(i) + "].color = RGBA(" + literal(*c) + ")");
            }
        } else if (t == "MATH") {
            inv.kernel = "Math<kMath::OPERATION_" + enum_token(param_str(node, "operation", "ADD")) + ", " +
                         (param_bool(node, "use_clamp") ? "true" : "false") + ">";
        } else if (t == "MIX_RGB") {
            inv.kernel = "ColorMix<kColorMix::BLEND_TYPE_" + enum_token(param_str(node, "blend_type", "MIX")) + ", " +
                         (param_bool(node, "use_clamp") ? "true" : "false") + ">";
        } else if (t == "BUMP") {
            inv.kernel = "Bump";
        } else if (t == "MIX_SHADER") {
            inv.kernel = "MixShader";
        } else if (t == "ADD_SHADER") {
            inv.kernel = "AddShader";
        } else if (t == "TEX_SKY_ANALYTIC") {
            inv.kernel = "TexSkyAnalytic";
        } else if (t == "LIGHT_PATH") {
            inv.kernel = "LightPath";
        } else if (t == "VECT_MATH") {
            const std::string op = enum_token(param_str(node, "operation", "ADD"));
            static const std::set<std::string> ops = {"ADD",     "SUBTRACT", "MULTIPLY", "MULTIPLY_ADD", "NORMALIZE",
                                                      "SCALE",   "MINIMUM",  "LENGTH",   "DOT_PRODUCT"};
            if (!ops.count(op)) throw CodegenError("unsupported vector math operation '" + op + "'");
            inv.kernel = "VectorMath<kVectorMath::OPERATION_" + op + ">";
        } else if (t == "TEX_CHECKER") {
            inv.kernel = "TexChecker";
        } else if (t == "SEPXYZ") {
            inv.kernel = "SeparateVector";
        } else if (t == "COMBXYZ") {
            inv.kernel = "CombineVector";
        } else if (t == "MAP_RANGE") {
            if (enum_token(param_str(node, "data_type", "FLOAT")) != "FLOAT" ||
                enum_token(param_str(node, "interpolation_type", "LINEAR")) != "LINEAR")
                throw CodegenError("map range: only float data with linear interpolation is supported");
            inv.kernel = std::string("MapRange<") + (param_bool(node, "clamp") ? "true" : "false") + ">";
        } else if (t == "MAPPING") {
            inv.kernel = "Mapping";
        } else if (t == "SEPARATE_COLOR" || t == "COMBINE_COLOR") {
            const std::string mode = enum_token(param_str(node, "mode", "RGB"));
            if (mode != "RGB" && mode != "HSV") throw CodegenError("unsupported color mode '" + mode + "'");
            inv.kernel = t == "SEPARATE_COLOR" ? "SeparateColor<kSeparateColor::MODE_" + mode + ">"
                                                : "CombineColor<kCombineColor::MODE_" + mode + ">";
        } else if (t == "HUE_SAT") {
            inv.kernel = "HueSaturation";
        } else if (t == "BRIGHTCONTRAST") {
            inv.kernel = "BrightnessContrast";
        } else if (t == "GAMMA") {
            inv.kernel = "Gamma";
        } else if (t == "INVERT") {
            inv.kernel = "ColorInvert";
        } else if (t == "RGBTOBW") {
            inv.kernel = "RGBToBW";
        } else if (t == "FRESNEL") {
            inv.kernel = "Fresnel";
        } else if (t == "LAYER_WEIGHT") {
            inv.kernel = "LayerWeight";
        } else if (t == "MIX") {
            inv.kernel = std::string("Mix<") + (param_bool(node, "clamp_factor") ? "true" : "false") + ", " +
                         (param_bool(node, "clamp_result") ? "true" : "false") + ", kMix::FACTOR_MODE_" +
                         enum_token(param_str(node, "factor_mode", "UNIFORM")) + ">";
        } else if (t == "TEX_GRADIENT") {
            const std::string gt = enum_token(param_str(node, "type", "LINEAR"));
            static const std::set<std::string> types = {"LINEAR", "SPHERICAL", "EASING", "QUADRATIC",
                                                        "QUADRATIC_SPHERE"};
            if (!types.count(gt)) throw CodegenError("unsupported gradient type '" + gt + "'");
            inv.kernel = "TexGradient<kTexGradient::TYPE_" + gt + ">";
        } else if (t == "BSDF_GLOSSY") {
            inv.kernel = "BsdfGlossy";
        } else if (t == "BSDF_ANISOTROPIC") {
            inv.kernel = "BsdfAnisotropic";
        } else if (t == "BSDF_GLASS") {
            inv.kernel = "BsdfGlass";
        } else if (t == "BSDF_REFRACTION") {
            inv.kernel = "BsdfRefraction";
        } else if (t == "BSDF_TRANSPARENT") {
            inv.kernel = "BsdfTransparent";
        } else if (t == "BSDF_TRANSLUCENT") {
            inv.kernel = "BsdfTranslucent";
        } else if (t == "BSDF_VELVET") {
            inv.kernel = "BsdfVelvet";
        } else if (t == "TEX_IMAGE" || t == "TEX_ENVIRONMENT") {
            inv = image_texture(node, t == "TEX_ENVIRONMENT");
        } else if (t == "TEX_SKY") {
            inv = sky_texture(node);
        } else if (t == "BSDF_HAIR") {
            inv.kernel = "BsdfHair";
        } else if (t == "TEX_WAVE") {
            // the reference's enums hold bands along the diagonal or x with a sine profile
            const std::string wt = enum_token(param_str(node, "type", "BANDS"));
            const std::string dir = enum_token(param_str(node, "direction", "X"));
            const std::string prof = enum_token(param_str(node, "profile", "SIN"));
            if (wt != "BANDS") throw CodegenError("unsupported wave type '" + wt + "'");
            if (dir != "DIAGONAL" && dir != "X") throw CodegenError("unsupported wave direction '" + dir + "'");
            if (prof != "SIN") throw CodegenError("unsupported wave profile '" + prof + "'");
            inv.kernel = "TexWave<kTexWave::TYPE_" + wt + ", kTexWave::DIRECTION_" + dir + ", kTexWave::PROFILE_" +
                         prof + ">";
        } else if (t == "NORMAL_MAP") {
            const std::string space = enum_token(param_str(node, "space", "TANGENT"));
            if (space != "TANGENT" && space != "WORLD") throw CodegenError("unsupported normal map space '" + space + "'");
            inv.kernel = "NormalMap<kNormalMap::SPACE_" + space + ">";
        } else if (const char* stub = stub_kernel(t)) {
            inv.kernel = stub;
        } else {
            throw CodegenError("unsupported node type '" + t + "'");
        }
        return inv;
    }

    // inputs whose exported default is meaningless (MaterialBuilder.swift:762-791)
    static void provide_defaults(const std::string& t, Invocation& inv, std::set<std::string>& cleared) {
        auto def = [&](const char* socket, const char* expr) {
            cleared.insert(socket);
            inv.in[make_identifier(socket)] = expr;
        };
        if (t == "TEX_IMAGE") def("Vector", "VECTOR(shading.uv)");
        else if (t == "TEX_CHECKER" || t == "TEX_NOISE" || t == "TEX_ENVIRONMENT" || t == "TEX_SKY_ANALYTIC")
            def("Vector", "VECTOR(shading.generated)");
        else if (t == "FRESNEL" || t == "LAYER_WEIGHT" || t == "BSDF_GLASS" || t == "BSDF_GLOSSY" ||
                 t == "BSDF_DIFFUSE" || t == "BSDF_PRINCIPLED" || t == "BSDF_TRANSLUCENT" || t == "BSDF_REFRACTION" ||
                 t == "BSDF_ANISOTROPIC" || t == "BSDF_VELVET" || t == "BUMP")
            def("Normal", "VECTOR(shading.normal)");
    }

    void emit_node(const Value& nodes, const std::string& key) {
        auto it = state_.find(key);
        if (it != state_.end()) {
            if (it->second == 2) return;
            throw CodegenError("cycle detected at node '" + key + "'");
        }
        const Value* node = nodes.get(key);
        if (!node) throw CodegenError("link to missing node '" + key + "'");
        state_[key] = 1;
        Invocation inv = make_invocation(*node);
        std::set<std::string> cleared;
        provide_defaults(node->get("type")->str, inv, cleared);
        if (const Value* inputs = node->get("inputs")) {
            std::vector<std::string> names;
            for (const auto& kv : inputs->obj) names.push_back(kv.first);
            std::sort(names.begin(), names.end());
            for (const auto& nm : names) {
                const Value& in = *inputs->get(nm);
                const Value* tv = in.get("type");
                const std::string type = coercion(tv && tv->is_string() ? tv->str : "VALUE");
                const Value* links = in.get("links");
                if (!links || !links->is_array() || links->arr.empty()) {
                    const Value* v = in.get("value");
                    if (v && !cleared.count(nm) && v->kind != Value::Null && v->kind != Value::String)
                        inv.in[make_identifier(nm)] = type + "(" + literal(*v) + ")";
                } else if (links->arr.size() == 1) {
                    const Value& l = links->arr[0];
                    std::string src, prop;
                    if (l.is_array() && l.arr.size() == 2) {
                        src = l.arr[0].str;
                        prop = l.arr[1].str;
                    } else if (l.is_object() && l.get("node") && l.get("property")) {
                        src = l.get("node")->str;
                        prop = l.get("property")->str;
                    } else {
                        throw CodegenError("malformed link on '" + key + "." + nm + "'");
                    }
                    emit_node(nodes, src);
                    inv.in[make_identifier(nm)] = type + "(n_" + make_identifier(src) + "." + make_identifier(prop) + ")";
                } else {
                    throw CodegenError("multi-input link on '" + key + "." + nm + "'");
                }
            }
        }
        invocations_.push_back({key, inv});
        state_[key] = 2;
    }
};

}  // namespace

std::string generate_material_source(const std::string& materials_json, bool noinline) {
    const json::Value root = json::parse(materials_json);
    Codegen cg;
    const json::Value* surface = root.get("surface");
    const json::Value* light = root.get("light");
    json::Value empty;
    empty.kind = json::Value::Array;
    cg.add_table("shade_surface", surface ? *surface : empty);
    cg.add_table("shade_light", light ? *light : empty);
    return cg.build(noinline);
}

std::string generate_material_textures(const std::string& materials_json) {
    const json::Value root = json::parse(materials_json);
    Codegen cg;
    const json::Value* surface = root.get("surface");
    const json::Value* light = root.get("light");
    json::Value empty;
    empty.kind = json::Value::Array;
    cg.add_table("shade_surface", surface ? *surface : empty);
    cg.add_table("shade_light", light ? *light : empty);
    std::string out = "[";
    for (size_t i = 0; i < cg.textures().size(); i++) {
        const TextureSlot& t = cg.textures()[i];
        char buf[256];
        snprintf(buf, sizeof(buf),
                 "\"sun_elevation\": %.9g, \"altitude\": %.9g, \"air_density\": %.9g, \"dust_density\": %.9g, "
                 "\"ozone_density\": %.9g",
                 t.sun_elevation, t.altitude, t.air_density, t.dust_density, t.ozone_density);
        out += std::string(i ? ", " : "") + "{\"slot\": " + std::to_string(i) + ", \"kind\": " + json_string(t.kind) +
               ", \"path\": " + json_string(t.path) + ", " + buf + "}";
    }
    return out + "]";
}

std::string material_identifier(const std::string& name) { return make_identifier(name); }

}  // namespace rm

extern "C" {

int rmh_codegen(const char* materials_json, int noinline, char** out_source, char** out_error) {
    if (out_source) *out_source = nullptr;
    if (out_error) *out_error = nullptr;
    if (!materials_json || !out_source) return -1;
    try {
        const std::string src = rm::generate_material_source(materials_json, noinline != 0);
        *out_source = strdup(src.c_str());
        return 0;
    } catch (const std::exception& e) {
        if (out_error) *out_error = strdup(e.what());
        return -1;
    }
}

int rmh_codegen_textures(const char* materials_json, char** out_json, char** out_error) {
    if (out_json) *out_json = nullptr;
    if (out_error) *out_error = nullptr;
    if (!materials_json || !out_json) return -1;
    try {
        *out_json = strdup(rm::generate_material_textures(materials_json).c_str());
        return 0;
    } catch (const std::exception& e) {
        if (out_error) *out_error = strdup(e.what());
        return -1;
    }
}

int rmh_identifier(const char* name, char* out, size_t capacity) {
    if (!name || !out || capacity == 0) return -1;
    const std::string id = rm::material_identifier(name);
    std::strncpy(out, id.c_str(), capacity - 1);
    out[capacity - 1] = '\0';
    return 0;
}

void rmh_free(void* p) { free(p); }

}  // extern "C"
