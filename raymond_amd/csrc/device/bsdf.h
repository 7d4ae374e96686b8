// Disney-style principled BSDF as a stochastic mixture of four lobes.
// Restates device/bsdf/{UberShader,BsdfSample,ShadingFrame,microfacet,
// fresnel}.hpp and device/bsdf/lobes/{Diffuse,Specular,Transmission,
// Clearcoat}.hpp.  Operation order follows the reference expressions
// left-to-right so the host and device builds agree bit for bit.
#pragma once
#include "common.h"
#include "elem.h"
#include "warp.h"

namespace rm {


// Shading-frame helpers: the frame normal is +z (device/bsdf/ShadingFrame.hpp:3-27)
namespace frame {
RM_HD bool same_hemisphere(vec3 wi, vec3 wo) { return wi.z * wo.z > 0.0f; }
RM_HD float cos_theta(vec3 w) { return w.z; }
RM_HD float cos_theta2(vec3 w) { return square(w.z); }
RM_HD float tan_theta2(vec3 w) { const float c2 = cos_theta2(w); return (1.0f - c2) / c2; }
}  // namespace frame

// (device/bsdf/BsdfSample.hpp:3-16)
struct BsdfSample {
    float pdf = 0.0f;
    vec3 wi;
    vec3 weight;
    uint32_t flags = 0;
    RM_HD static BsdfSample invalid() {
        BsdfSample r;
        r.pdf = 0.0f;
        r.weight = vec3(0.0f);
        r.flags = 0;
        return r;
    }
};

// ---------------------------------------------------------------------------
// Fresnel (device/bsdf/fresnel.hpp:3-64)

RM_HD float schlick_weight(float cos_theta) {
    const float m = saturate(1.0f - cos_theta);
    return ((m * m) * (m * m)) * m;
}

RM_HD float fresnel_dielectric_cos(float cosi, float eta) {
    const float c = abs_(cosi);
    float g = (eta * eta - 1.0f) + c * c;
    if (g > 0.0f) {
        g = sqrt_(g);
        const float A = (g - c) / (g + c);
        const float B = (c * (g + c) - 1.0f) / (c * (g - c) + 1.0f);
        return ((0.5f * A) * A) * (1.0f + B * B);
    }
    return 1.0f;
}

RM_HD vec3 interpolate_fresnel(vec3 wi, vec3 wh, float ior, float F0, vec3 cspec0) {
    const float F0_norm = 1.0f / (1.0f - F0);
    const float FH = (fresnel_dielectric_cos(dot(wi, wh), ior) - F0) * F0_norm;
    return cspec0 * (1.0f - FH) + vec3(1.0f) * FH;
}

RM_HD vec3 fresnel_reflection_color(vec3 wi, vec3 wh, float ior, vec3 cspec0) {
    const float F0 = fresnel_dielectric_cos(1.0f, ior);
    return interpolate_fresnel(wi, wh, ior, F0, cspec0);
}

// ---------------------------------------------------------------------------
// Microfacet distributions (device/bsdf/microfacet.hpp:8-108)

// isotropic GTR1 (Burley 2012)
RM_HD float gtr1(vec3 wh, float a) {
    const float n_dot_h = frame::cos_theta(wh);
    const float a2 = square(a);
    const float t = 1.0f + (a2 - 1.0f) * square(n_dot_h);
    return (a2 - 1.0f) / ((kPi * log_(a2)) * t);
}

RM_HD vec3 sample_gtr1(vec2 rnd, float a) {
    const float a2 = square(a);
    const float cos_theta = safe_sqrt((1.0f - pow_(a2, 1.0f - rnd.x)) / (1.0f - a2));
    const float sin_theta = safe_sqrt(1.0f - (cos_theta * cos_theta));
    const float phi = 2.0f * kPi * rnd.y;
    const float sin_phi = sin_(phi);
    const float cos_phi = cos_(phi);
    return vec3(sin_theta * cos_phi, sin_theta * sin_phi, cos_theta);
}

RM_HD float smith_g1(vec3 w, vec3 wh, float a) {
    if ((dot(w, wh) * frame::cos_theta(w)) * frame::cos_theta(wh) <= 0.0f) return 0.0f;
    if (abs_(frame::cos_theta(w)) >= 1.0f) return 1.0f;
    const float a2tan2 = square(a) * frame::tan_theta2(w);
    return 2.0f / (1.0f + sqrt_(1.0f + a2tan2));
}

RM_HD float aniso_smith_g1(vec3 w, vec3 wh, float ax, float ay) {
    if ((dot(w, wh) * frame::cos_theta(w)) * frame::cos_theta(wh) <= 0.0f) return 0.0f;
    if (abs_(frame::cos_theta(w)) >= 1.0f) return 1.0f;
    const float a2tan2 = (square(ax * w.x) + square(ay * w.y)) / frame::cos_theta2(w);
    return 2.0f / (1.0f + sqrt_(1.0f + a2tan2));
}

// anisotropic GGX D (Walter et al. 2007)
RM_HD float aniso_ggx(vec3 wh, float ax, float ay) {
    const float n_dot_h = frame::cos_theta(wh);
    const float a = wh.x / ax;
    const float b = wh.y / ay;
    const float c = (square(a) + square(b)) + square(n_dot_h);
    return 1.0f / (((kPi * ax) * ay) * square(c));
}

// GGX visible-normal sampling (Heitz 2018), with the sign flip for wo below
// the surface (device/bsdf/microfacet.hpp:87-108)
RM_HD vec3 sample_ggx_vndf(vec2 rnd, float ax, float ay, vec3 wo) {
    const float sgn = sign_(frame::cos_theta(wo));
    const vec3 Vh = sgn * normalize(vec3(ax * wo.x, ay * wo.y, wo.z));
    const float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    const vec3 T1 = lensq > 0.0f ? vec3(-Vh.y, Vh.x, 0.0f) * (1.0f / sqrt_(lensq)) : vec3(1.0f, 0.0f, 0.0f);
    const vec3 T2 = cross(Vh, T1);
    const float r = sqrt_(rnd.x);
    const float phi = 2.0f * kPi * rnd.y;
    const float t1 = r * cos_(phi);
    float t2 = r * sin_(phi);
    const float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * sqrt_(1.0f - t1 * t1) + s * t2;
    const vec3 Nh = (t1 * T1 + t2 * T2) + sqrt_(max_(0.0f, (1.0f - t1 * t1) - t2 * t2)) * Vh;
    const vec3 Ne = normalize(vec3(ax * Nh.x, ay * Nh.y, max_(0.0f, Nh.z)));
    return sgn * Ne;
}

// ---------------------------------------------------------------------------
// Lobes

// Burley diffuse with retro-reflection and sheen (lobes/Diffuse.hpp:14-96)
struct Diffuse {
    vec3 diffuse_weight;
    vec3 sheen_weight;
    float roughness = 0.0f;
    bool translucent = false;

    RM_HD vec3 shade(vec3 wo, vec3 wi, float NdotL) const {
        const float NdotV = abs_(frame::cos_theta(wo));
        const float LdotV = dot(wi, wo);
        const float FL = schlick_weight(NdotL);
        const float FV = schlick_weight(NdotV);
        const float lambertian = (1.0f - 0.5f * FV) * (1.0f - 0.5f * FL);
        const float LH2 = LdotV + 1.0f;
        const float RR = roughness * LH2;
        const float retro = RR * ((FL + FV) + (FL * FV) * (RR - 1.0f));
        const vec3 wh = normalize(wo + wi);
        const float LdotH = abs_(dot(wh, wi));
        const float sheen = schlick_weight(LdotH);
        return diffuse_weight * (lambertian + retro) + sheen_weight * (kPi * sheen);
    }

    RM_HD vec3 evaluate(vec3 wo, vec3 wi, float& pdf) const {
        if (frame::same_hemisphere(wi, wo) == translucent) {
            pdf = 0.0f;
            return vec3(0.0f);
        }
        const float NdotL = abs_(frame::cos_theta(wi));
        pdf = kInvPi * NdotL;
        return pdf * shade(wo, wi, NdotL);
    }

    RM_HD BsdfSample sample(vec2 rnd, vec3 wo) const {
        BsdfSample r;
        r.wi = warp::uniform_square_to_cosine_hemisphere(rnd);
        if (!frame::same_hemisphere(r.wi, wo)) r.wi = r.wi * -1.0f;
        const float NdotL = abs_(frame::cos_theta(r.wi));
        r.pdf = kInvPi * NdotL;
        if (!(r.pdf > 0.0f)) return BsdfSample::invalid();
        r.weight = shade(wo, r.wi, NdotL);
        if (translucent) {
            r.wi = -r.wi;
            r.flags = RayTransmission | RayDiffuse;
        } else {
            r.flags = RayReflection | RayDiffuse;
        }
        return r;
    }
};

// Anisotropic GGX reflection (lobes/Specular.hpp:15-67).  evaluate() omits
// `weight` while sample() includes it, as in the reference.
struct Specular {
    float alpha_x = 0.0f;
    float alpha_y = 0.0f;
    vec3 cspec0;
    float ior = 0.0f;
    float weight = 0.0f;

    RM_HD vec3 evaluate(vec3 wo, vec3 wi, float& pdf) const {
        const vec3 wh = normalize(wi + wo);
        pdf = (aniso_ggx(wh, alpha_x, alpha_y) * aniso_smith_g1(wo, wh, alpha_x, alpha_y)) *
              abs_(dot(wo, wh) / frame::cos_theta(wo));
        if (!(pdf > 0.0f)) {
            pdf = 0.0f;
            return vec3(0.0f);
        }
        pdf = pdf * (1.0f / abs_(4.0f * dot(wo, wh)));
        const vec3 F = fresnel_reflection_color(wi, wh, ior, cspec0);
        const float G = aniso_smith_g1(wi, wh, alpha_x, alpha_y) * aniso_smith_g1(wo, wh, alpha_x, alpha_y);
        const float D = aniso_ggx(wh, alpha_x, alpha_y);
        return ((F * D) * G) / abs_(4.0f * frame::cos_theta(wo));
    }

    RM_HD BsdfSample sample(vec2 rnd, vec3 wo) const {
        BsdfSample r;
        const vec3 wh = sample_ggx_vndf(rnd, alpha_x, alpha_y, wo);
        r.pdf = (aniso_ggx(wh, alpha_x, alpha_y) * aniso_smith_g1(wo, wh, alpha_x, alpha_y)) *
                abs_(dot(wo, wh) / frame::cos_theta(wo));
        if (!(r.pdf > 0.0f)) return BsdfSample::invalid();
        r.wi = reflect(-wo, wh);
        if (!frame::same_hemisphere(r.wi, wo)) return BsdfSample::invalid();
        r.pdf = r.pdf * (1.0f / abs_(4.0f * dot(wo, wh)));
        const vec3 F = fresnel_reflection_color(r.wi, wh, ior, cspec0);
        const float Gi = aniso_smith_g1(r.wi, wh, alpha_x, alpha_y);
        r.weight = (weight * F) * Gi;
        r.flags = RayReflection | RayGlossy;
        return r;
    }
};

// Rough dielectric reflection + refraction (lobes/Transmission.hpp:15-123).
struct Transmission {
    float reflection_alpha = 0.0f;
    float transmission_alpha = 0.0f;
    vec3 base_color;
    vec3 cspec0;
    float ior = 0.0f;
    float weight = 0.0f;
    bool only_refract = false;

    RM_HD vec3 evaluate(vec3 wo, vec3 wi, float& pdf) const {
        const bool is_reflection = frame::same_hemisphere(wi, wo);
        if (only_refract && is_reflection) {
            pdf = 0.0f;
            return vec3(0.0f);
        }
        const float eta = frame::cos_theta(wo) > 0.0f ? ior : 1.0f / ior;
        const vec3 wh = is_reflection ? normalize(wi + wo) : normalize(wi * eta + wo);
        const float alpha = select(is_reflection, reflection_alpha, transmission_alpha);
        pdf = (aniso_ggx(wh, alpha, alpha) * aniso_smith_g1(wo, wh, alpha, alpha)) *
              abs_(dot(wo, wh) / frame::cos_theta(wo));
        if (!(pdf > 0.0f)) {
            pdf = 0.0f;
            return vec3(0.0f);
        }
        const float Gi = aniso_smith_g1(wi, wh, alpha, alpha);
        const float Fr = only_refract ? 0.0f : fresnel_dielectric_cos(frame::cos_theta(wo), eta);
        if (is_reflection) {
            pdf = pdf * Fr;
            pdf = pdf * (1.0f / abs_(4.0f * dot(wo, wh)));
            const vec3 F = fresnel_reflection_color(wi, wh, eta, cspec0);
            return ((pdf * weight) * F) * Gi;
        } else {
            pdf = pdf * (1.0f - Fr);
            pdf = pdf * abs_(dot(wi, wh) / square(dot(wi, wh) + dot(wh, wo) / eta));
            return ((pdf * weight) * base_color) * Gi;
        }
    }

    RM_HD BsdfSample sample(vec2 rnd, vec3 wo) const {
        BsdfSample r;
        const float eta = frame::cos_theta(wo) > 0.0f ? ior : 1.0f / ior;
        const float Fr = only_refract ? 0.0f : fresnel_dielectric_cos(frame::cos_theta(wo), eta);
        const bool is_reflection = rnd.x < Fr;
        const float alpha = select(is_reflection, reflection_alpha, transmission_alpha);
        // VNDF sampling consumes the unscaled rnd (lobes/Transmission.hpp:71-84)
        const vec3 wh = sample_ggx_vndf(rnd, alpha, alpha, wo);
        r.pdf = (aniso_ggx(wh, alpha, alpha) * aniso_smith_g1(wo, wh, alpha, alpha)) *
                abs_(dot(wo, wh) / frame::cos_theta(wo));
        if (is_reflection) {
            if (!(r.pdf > 0.0f)) return BsdfSample::invalid();
            r.wi = reflect(-wo, wh);
            if (!frame::same_hemisphere(r.wi, wo)) return BsdfSample::invalid();
            r.pdf = r.pdf * Fr;
            r.pdf = r.pdf * (1.0f / abs_(4.0f * dot(wo, wh)));
            const vec3 F = fresnel_reflection_color(r.wi, wh, eta, cspec0);
            const float Gi = aniso_smith_g1(r.wi, wh, alpha, alpha);
            r.weight = (weight * F) * Gi;
            r.flags = RayReflection | RayGlossy;
            return r;
        } else {
            if (!(r.pdf > 0.0f)) return BsdfSample::invalid();
            r.wi = refract(-wo, wh, 1.0f / eta);
            if (frame::same_hemisphere(r.wi, wo)) return BsdfSample::invalid();
            r.pdf = r.pdf * (1.0f - Fr);
            r.pdf = r.pdf * abs_(dot(r.wi, wh) / square(dot(r.wi, wh) + dot(wh, wo) / eta));
            const float Gi = aniso_smith_g1(r.wi, wh, alpha, alpha);
            r.weight = (weight * base_color) * Gi;
            r.flags = RayTransmission | RayGlossy;
            return r;
        }
    }
};

// Clearcoat: GGX-sampled, Smith-shadowed, fixed IOR 1.5 (lobes/Clearcoat.hpp:15-63)
struct Clearcoat {
    float alpha = 0.0f;
    float weight = 0.0f;

    RM_HD vec3 evaluate(vec3 wo, vec3 wi, float& pdf) const {
        const vec3 wh = normalize(wi + wo);
        pdf = (aniso_ggx(wh, alpha, alpha) * aniso_smith_g1(wo, wh, alpha, alpha)) *
              abs_(dot(wo, wh) / frame::cos_theta(wo));
        if (!(pdf > 0.0f)) {
            pdf = 0.0f;
            return vec3(0.0f);
        }
        pdf = pdf * (1.0f / abs_(4.0f * dot(wo, wh)));
        const vec3 F = fresnel_reflection_color(wi, wh, 1.5f, vec3(0.04f));
        const float G = aniso_smith_g1(wi, wh, alpha, alpha) * aniso_smith_g1(wo, wh, alpha, alpha);
        const float D = aniso_ggx(wh, alpha, alpha);
        return (((0.25f * F) * D) * G) / abs_(4.0f * frame::cos_theta(wo));
    }

    RM_HD BsdfSample sample(vec2 rnd, vec3 wo) const {
        BsdfSample r;
        const vec3 wh = sample_ggx_vndf(rnd, alpha, alpha, wo);
        r.pdf = (aniso_ggx(wh, alpha, alpha) * aniso_smith_g1(wo, wh, alpha, alpha)) *
                abs_(dot(wo, wh) / frame::cos_theta(wo));
        if (!(r.pdf > 0.0f)) return BsdfSample::invalid();
        r.wi = reflect(-wo, wh);
        if (!frame::same_hemisphere(r.wi, wo)) return BsdfSample::invalid();
        r.pdf = r.pdf * (1.0f / abs_(4.0f * dot(wo, wh)));
        const vec3 F = fresnel_reflection_color(r.wi, wh, 1.5f, vec3(0.04f));
        const float Gi = smith_g1(r.wi, wh, alpha);
        r.weight = ((0.25f * weight) * F) * Gi;
        r.flags = RayReflection | RayGlossy;
        return r;
    }
};

// ---------------------------------------------------------------------------
// The mixture (device/bsdf/UberShader.hpp:10-146)

struct UberShader {
    vec3 normal;
    float lobe_prob[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    Diffuse diffuse;
    Specular specular;
    Transmission transmission;
    Clearcoat clearcoat;
    float alpha = 1.0f;
    vec3 alpha_weight = vec3(1.0f);
    vec3 emission = vec3(0.0f);
    float weight = 1.0f;

    RM_HD bool is_delta() const {
        if (alpha < 0.5f) return true;
        if (specular.weight > 0.5f && (specular.alpha_x < 0.1f || specular.alpha_y < 0.1f)) return true;
        if (transmission.weight > 0.5f &&
            (transmission.reflection_alpha < 0.1f || transmission.transmission_alpha < 0.1f))
            return true;
        return false;
    }

    RM_HD vec3 albedo() const {
        vec3 v(0.0f);
        if (lobe_prob[0] > 0.0f) v += diffuse.diffuse_weight + diffuse.sheen_weight;
        if (lobe_prob[1] > 0.0f) v += (specular.weight * (specular.cspec0 + 1.0f)) / 2.0f;
        if (lobe_prob[2] > 0.0f)
            v += (transmission.weight * ((transmission.cspec0 + transmission.base_color) + 2.0f)) / 4.0f;
        if (lobe_prob[3] > 0.0f) v += vec3(clearcoat.weight / 4.0f);
        v = alpha * v + (1.0f - alpha) * alpha_weight;
        return v + 1e-3f;
    }

    // Mixture value and pdf in the shading frame.
    RM_HD vec3 evaluate_local(vec3 wo, vec3 wi, float& pdf) const {
        pdf = 0.0f;
        vec3 value(0.0f);
        float lp;
        if (lobe_prob[0] > 0.0f) { value += diffuse.evaluate(wo, wi, lp); pdf += lobe_prob[0] * lp; }
        if (lobe_prob[1] > 0.0f) { value += specular.evaluate(wo, wi, lp); pdf += lobe_prob[1] * lp; }
        if (lobe_prob[2] > 0.0f) { value += transmission.evaluate(wo, wi, lp); pdf += lobe_prob[2] * lp; }
        if (lobe_prob[3] > 0.0f) { value += clearcoat.evaluate(wo, wi, lp); pdf += lobe_prob[3] * lp; }
        pdf = pdf * alpha;
        value = value * (weight * alpha);
        return value;
    }

    // World-space evaluation with the shading/geometric hemisphere checks.
    RM_HD vec3 evaluate(vec3 wo, vec3 wi, vec3 sh_normal, vec3 geo_normal, float& pdf) const {
        pdf = 0.0f;
        const mat3 to_local = build_onb(sh_normal);
        const float wo_geo = dot(wo, geo_normal);
        wo = wo * to_local;
        if (wo.z * wo_geo < 0.0f) return vec3(0.0f);
        const float wi_geo = dot(wi, geo_normal);
        wi = wi * to_local;
        if (wi.z * wi_geo < 0.0f) return vec3(0.0f);
        return evaluate_local(wo, wi, pdf);
    }

    RM_HD BsdfSample sample(vec3 rnd, vec3 wo, vec3 sh_normal, vec3 geo_normal, uint32_t previous_flags) const {
        if (rnd.x < alpha) {
            rnd.x = rnd.x / alpha;
        } else {
            // null scattering keeps the ray flags (UberShader.hpp:86-96)
            BsdfSample s;
            s.weight = alpha_weight * weight;
            s.wi = -wo;
            s.pdf = 1.0f;
            s.flags = previous_flags | RaySingular;
            return s;
        }
        const mat3 to_local = build_onb(sh_normal);
        const float wo_geo = dot(wo, geo_normal);
        wo = wo * to_local;
        if (wo.z * wo_geo < 0.0f) return BsdfSample::invalid();

        // lobe probabilities as values, not addresses: keeps the mixture in registers
        const float lp0 = lobe_prob[0], lp1 = lobe_prob[1], lp2 = lobe_prob[2], lp3 = lobe_prob[3];
        BsdfSample s;
        float p_lobe;
        const vec2 r2(rnd.y, rnd.z);
        if (rnd.x < lp0) {
            p_lobe = lp0; s = diffuse.sample(r2, wo);
        } else if (rnd.x < (lp0 + lp1)) {
            p_lobe = lp1; s = specular.sample(r2, wo);
        } else if (rnd.x < ((lp0 + lp1) + lp2)) {
            p_lobe = lp2; s = transmission.sample(r2, wo);
        } else {
            p_lobe = lp3; s = clearcoat.sample(r2, wo);
        }
        if (!isfinite_(p_lobe)) return BsdfSample::invalid();
        if (p_lobe < 1.0f) {
            // MIS needs the whole mixture, not just the sampled lobe (UberShader.hpp:127-131)
            s.weight = evaluate_local(wo, s.wi, s.pdf);
            s.weight = s.weight / s.pdf;
        } else {
            s.pdf = s.pdf * alpha;
            s.weight = s.weight * (weight * alpha);
        }
        const float wi_sh = s.wi.z;
        s.wi = to_local * s.wi;
        const float wi_geo = dot(s.wi, geo_normal);
        if (wi_sh * wi_geo < 0.0f) return BsdfSample::invalid();
        return s;
    }
};

}  // namespace rm
