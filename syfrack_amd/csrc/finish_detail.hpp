// Finish detail patterns (finish_detail.wgsl; applied by make_surface, pathtrace.wgsl:306-329):
// world-space procedural relief for a named finish.  One evaluation returns the height w in
// [0, 1] and its analytic slope xyz; the height blends the finish toward its tint and moves its
// roughness, the slope tilts the shading normal by the bump amount.  Every band fades to its
// mean where the ray footprint (the path's texture width, in scene units) grows past its
// wavelength, so distant relief never aliases.  Patterns 1..12: VeinedStone, Patina,
// BrushedMetal, Basalt, Ripples, Rosettes, CoastalRock, Sand, Dust, Concrete, WeatheredConcrete,
// WetConcrete.  This path has no dunes (the sand's wind direction is zero) and no ocean (the
// coastal rock's tide line is inert).
//
// Included by shade.hpp ahead of make_surface, after the vector and mixing helpers it uses.
#pragma once

namespace syf {

struct F4 {
    F3 g;     // slope
    float w;  // height
};

__device__ __forceinline__ float smoothstepf(float lo, float hi, float x) {
    const float t = clampT((x - lo) / (hi - lo), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
// d/dx smoothstep(lo, hi, x) (finish_smooth_derivative)
__device__ __forceinline__ float smoothstep_slope(float lo, float hi, float x) {
    const float t = clampT((x - lo) / (hi - lo), 0.0f, 1.0f);
    return 6.0f * t * (1.0f - t) / (hi - lo);
}
__device__ __forceinline__ float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ float fractf(float x) { return x - floorf(x); }

// A cell's value in [0, 1): 24 bits of a PCG hash of the cell's integer coordinates and the seed.
__device__ __forceinline__ float cell_value(int cx, int cy, int cz, uint32_t seed) {
    const uint32_t h = pcg(uint32_t(cx) ^ pcg(uint32_t(cy) + 0x9e3779b9u) ^ pcg(uint32_t(cz) + seed));
    return float(h >> 8) * (1.0f / 16777216.0f);
}

// Value noise with C1 (smoothstep) trilinear weights and its exact gradient.
__device__ inline F4 value_noise(F3 p, uint32_t seed) {
    const float fx = floorf(p.x), fy = floorf(p.y), fz = floorf(p.z);
    const int cx = int(fx), cy = int(fy), cz = int(fz);
    const float ux = p.x - fx, uy = p.y - fy, uz = p.z - fz;
    const float tx = ux * ux * (3.0f - 2.0f * ux), ty = uy * uy * (3.0f - 2.0f * uy), tz = uz * uz * (3.0f - 2.0f * uz);
    const float sx = 6.0f * ux * (1.0f - ux), sy = 6.0f * uy * (1.0f - uy), sz = 6.0f * uz * (1.0f - uz);
    const float v000 = cell_value(cx, cy, cz, seed), v100 = cell_value(cx + 1, cy, cz, seed);
    const float v010 = cell_value(cx, cy + 1, cz, seed), v110 = cell_value(cx + 1, cy + 1, cz, seed);
    const float v001 = cell_value(cx, cy, cz + 1, seed), v101 = cell_value(cx + 1, cy, cz + 1, seed);
    const float v011 = cell_value(cx, cy + 1, cz + 1, seed), v111 = cell_value(cx + 1, cy + 1, cz + 1, seed);
    const float a = mixf(v000, v100, tx), b = mixf(v010, v110, tx);
    const float c = mixf(v001, v101, tx), d = mixf(v011, v111, tx);
    const float e = mixf(a, b, ty), g = mixf(c, d, ty);
    const float dx = mixf(mixf(v100 - v000, v110 - v010, ty), mixf(v101 - v001, v111 - v011, ty), tz);
    return F4{mk(dx * sx, mixf(b - a, d - c, tz) * sy, (g - e) * sz), mixf(e, g, tz)};
}

// Three octaves (frequency x2.13, weight /2), each fading to its mean 0.5 as the footprint
// `width` reaches its wavelength; normalised by the weights' sum 0.875.
__device__ inline F4 octaves(F3 p, float width, uint32_t seed) {
    float frequency = 1.0f, weight = 0.5f;
    F4 r{mk(0.0f, 0.0f, 0.0f), 0.0f};
    for (uint32_t i = 0; i < 3u; ++i) {
        const float visible = 1.0f - smoothstepf(0.35f, 0.85f, width * frequency);
        if (visible > 0.0f) {
            const F4 n = value_noise(scl(p, frequency), seed + i * 97u);
            r.g = add(r.g, scl(n.g, weight * frequency * visible));
            r.w += weight * mixf(0.5f, n.w, visible);
        } else {
            r.w += weight * 0.5f;
        }
        frequency *= 2.13f;
        weight *= 0.5f;
    }
    return F4{mk(r.g.x / 0.875f, r.g.y / 0.875f, r.g.z / 0.875f), r.w / 0.875f};
}
__device__ __forceinline__ F3 mul3c(F3 a, float x, float y, float z) { return mk(a.x * x, a.y * y, a.z * z); }

// finish_eval: the pattern's height and slope at world point p for footprint `footprint`.
__device__ __attribute__((noinline)) F4 finish_pattern(uint32_t kind, uint32_t seed, float scale, F3 p, float footprint) {
    const F3 q = scl(p, scale);
    const float width = footprint * scale;
    if (kind == 12u) {  // WetConcrete: rain pools in broad basins; the pools' ripples replace the grit
        const F4 basin = octaves(mul3c(q, 0.31f, 0.31f, 0.22f), width * 0.31f, seed + 911u);
        const float wet = smoothstepf(0.47f, 0.59f, basin.w);
        const F4 grit = octaves(scl(q, 16.0f), width * 16.0f, seed + 113u);
        const float visible = 1.0f - smoothstepf(0.15f, 0.6f, width * 3.0f);
        const float phase = q.x * 2.8f + q.z * 4.1f;
        const F3 ripple = scl(mk(2.8f, 0.0f, 4.1f), cosf(phase) * 0.018f * visible);
        return F4{scl(mix3(scl(grit.g, 16.0f), ripple, wet), scale), wet};
    }
    if (kind == 10u || kind == 11u) {  // Concrete: aggregate, pores, stains and board joints
        const F4 stains = octaves(scl(q, 0.14f), width * 0.14f, seed + 17u);
        const F4 aggregate = octaves(scl(q, 7.0f), width * 7.0f, seed + 113u);
        const F4 pore = value_noise(scl(q, 23.0f), seed + 307u);
        const float visible = 1.0f - smoothstepf(0.25f, 0.75f, width * 23.0f);
        const float pit = 1.0f - smoothstepf(0.24f, 0.43f, pore.w);
        const F3 pit_slope = scl(pore.g, -smoothstep_slope(0.24f, 0.43f, pore.w) * 23.0f * visible);
        const float phase = fractf(q.y / 1.25f + 0.5f) - 0.5f;
        const float board_visible = 1.0f - smoothstepf(0.03f, 0.15f, width);
        const float joint = (1.0f - smoothstepf(0.006f, 0.027f, fabsf(phase))) * board_visible;
        const float joint_slope = -smoothstep_slope(0.006f, 0.027f, fabsf(phase)) * signf(phase) / 1.25f * board_visible;
        const float height = 0.18f * stains.w + 0.65f * aggregate.w - 0.14f * mixf(0.2f, pit, visible) - 0.09f * joint;
        F3 slope = add(scl(stains.g, 0.18f * 0.14f), scl(aggregate.g, 0.65f * 7.0f));
        slope = sub(slope, scl(pit_slope, 0.14f));
        slope = scl(sub(slope, mk(0.0f, 0.09f * joint_slope, 0.0f)), scale);
        if (kind == 11u) {  // WeatheredConcrete: vertical rain streaks deposit over the stains
            const F4 rain = octaves(mul3c(q, 1.9f, 0.055f, 1.9f), width * 1.9f, seed + 601u);
            const float deposit = smoothstepf(0.33f, 0.68f, rain.w);
            const F3 rain_slope = scl(mul3c(rain.g, 1.9f, 0.055f, 1.9f), scale * smoothstep_slope(0.33f, 0.68f, rain.w));
            return F4{add(slope, scl(rain_slope, 0.12f)), clampT(0.72f * deposit + 0.28f * stains.w, 0.0f, 1.0f)};
        }
        return F4{slope, clampT(height + 0.12f, 0.0f, 1.0f)};
    }
    if (kind == 9u) {  // Dust: domain-warped deposits with powder grain; slopes through the warp's Jacobian
        const F4 wa = octaves(scl(q, 0.17f), width * 0.17f, seed + 17u);
        const F4 wb = octaves(scl(q, 0.17f), width * 0.17f, seed + 53u);
        const F3 warped = add(q, mk(2.8f * (wa.w - 0.5f), 0.0f, 2.8f * (wb.w - 0.5f)));
        const F4 cloud = octaves(scl(warped, 0.16f), width * 0.48f, seed + 211u);
        const F4 streak = octaves(mul3c(warped, 0.65f, 0.18f, 0.09f), width * 1.95f, seed + 397u);
        const float density = 0.78f * cloud.w + 0.22f * streak.w;
        const F3 g = add(scl(cloud.g, 0.78f * 0.16f), scl(mul3c(streak.g, 0.65f, 0.18f, 0.09f), 0.22f));
        const F3 density_slope = scl(add(g, scl(add(scl(wa.g, g.x), scl(wb.g, g.z)), 2.8f * 0.17f)), scale);
        const float deposit = smoothstepf(0.36f, 0.64f, density);
        const F3 deposit_slope = scl(density_slope, smoothstep_slope(0.36f, 0.64f, density));
        const F4 grain = octaves(scl(p, 48.0f), footprint * 48.0f, seed + 739u);
        const float powder = 0.85f + 0.15f * grain.w;
        return F4{add(scl(deposit_slope, powder), scl(grain.g, deposit * 0.15f * 48.0f)), deposit * powder};
    }
    if (kind == 8u) {  // Sand: wind ripples (wind along +x) over fine grain
        const float xx = p.x * scale, zz = p.z * scale;
        const F4 warp = value_noise(mk(xx * 0.045f, 0.0f, zz * 0.045f), seed);
        const float phase = 2.0f * PI_F * (xx + 1.7f * warp.w + 0.12f * sinf(zz * 0.22f));
        const float wv = width * 3.0f;
        const float visible = expf(-(wv * wv));
        const float ripple = 0.5f + 0.36f * sinf(phase) + 0.10f * sinf(2.0f * phase);
        const float k = 2.0f * PI_F * (0.36f * cosf(phase) + 0.20f * cosf(2.0f * phase)) * scale * visible;
        const float gx = k * (1.0f + 0.0765f * warp.g.x);
        const float gz = k * (0.0264f * cosf(zz * 0.22f) + 0.0765f * warp.g.z);
        const F4 grain = octaves(scl(p, 90.0f), footprint * 90.0f, seed + 139u);
        return F4{add(mk(gx, 0.0f, gz), scl(grain.g, 2.5f)), 0.70f * mixf(0.5f, ripple, visible) + 0.30f * grain.w};
    }
    if (kind == 6u) {  // Rosettes: carved eight-petal flowers in rings on a unit grid in xz
        const float tx = fractf(q.x + 0.5f) - 0.5f, tz = fractf(q.z + 0.5f) - 0.5f;
        const float r = fmaxf(sqrtf(tx * tx + tz * tz), 1e-5f);
        const float angle = atan2f(tz, tx);
        const float flower = r - 0.25f - 0.07f * cosf(angle * 8.0f);
        const float ring = r - 0.45f;
        const float ax = -tz / (r * r), az = tx / (r * r);
        const float sf = signf(flower), s8 = 0.56f * sinf(angle * 8.0f);
        float g2x = sf * (tx / r + s8 * ax), g2z = sf * (tz / r + s8 * az);
        float distance = fabsf(flower);
        if (fabsf(ring) < distance) {
            distance = fabsf(ring);
            g2x = signf(ring) * tx / r;
            g2z = signf(ring) * tz / r;
        }
        const float visible = 1.0f - smoothstepf(0.10f, 0.40f, width);
        const float carving = 1.0f - smoothstepf(0.012f, 0.045f, distance);
        const float k = -smoothstep_slope(0.012f, 0.045f, distance) * visible * scale;
        return F4{mk(g2x * k, 0.0f, g2z * k), mixf(0.2f, carving, visible)};
    }
    if (kind == 3u) {  // BrushedMetal: noise stretched across the brushing direction
        const F4 n = octaves(mul3c(q, 18.0f, 0.45f, 3.0f), width * 18.0f, seed);
        return F4{scl(mul3c(n.g, 18.0f, 0.45f, 3.0f), scale), n.w};
    }
    const F4 n = octaves(q, width, seed);
    F4 r = n;
    if (kind == 1u) {  // VeinedStone: turbulent sine veins
        const F3 axis = mk(0.16f, 0.11f, 0.08f);
        const float phase = (dot3(q, axis) + n.w * 2.8f) * PI_F * 2.0f;
        const float wave = sinf(phase);
        const float veins = 1.0f - smoothstepf(0.035f, 0.19f, fabsf(wave));
        const F3 g = scl(add(axis, scl(n.g, 2.8f)),
                         -smoothstep_slope(0.035f, 0.19f, fabsf(wave)) * signf(wave) * cosf(phase) * PI_F * 2.0f);
        const float visible = 1.0f - smoothstepf(0.18f, 0.45f, width);
        const F3 vg = add(scl(n.g, 0.65f), scl(g, 0.35f));
        const float vw = 0.65f * n.w + 0.35f * veins;
        r = F4{mix3(n.g, vg, visible), mixf(n.w, vw, visible)};
    } else if (kind == 2u) {  // Patina: thresholded blotches
        r = F4{scl(n.g, smoothstep_slope(0.35f, 0.67f, n.w)), smoothstepf(0.35f, 0.67f, n.w)};
    } else if (kind == 4u || kind == 7u) {  // Basalt, CoastalRock: noise over finer grain
        const F4 grain = octaves(scl(q, 3.7f), width * 3.7f, seed + 53u);
        r = F4{add(scl(n.g, 0.8f), scl(grain.g, 0.2f * 3.7f)), 0.8f * n.w + 0.2f * grain.w};
    } else if (kind == 5u) {  // Ripples: two crossing sine trains bent by the noise
        const float a = q.x + 0.75f * n.w, b = q.z * 1.37f + q.x * 0.4f;
        const float visible = 1.0f - smoothstepf(0.35f, 1.0f, width);
        const F3 g = add(scl(add(mk(1.0f, 0.0f, 0.0f), scl(n.g, 0.75f)), 0.32f * cosf(a)),
                         scl(mk(0.4f, 0.0f, 1.37f), 0.18f * cosf(b)));
        r = F4{scl(g, visible), 0.5f + (0.32f * sinf(a) + 0.18f * sinf(b)) * visible};
    }
    return F4{scl(r.g, scale), r.w};
}

// The finish's detail applied to the surface parameters (pathtrace.wgsl:306-329).
__device__ __forceinline__ void apply_finish_detail(const FinishP& f, F3 p, F3 n, float footprint, F3& albedo,
                                                    float& rough, float& metallic, float& spec, F3& shading_n) {
    F4 d = finish_pattern(f.pattern, f.seed, f.scale, p, footprint);
    if (f.pattern == 9u) {  // settled powder gathers on tops and recedes on steep bevels
        const float up = smoothstepf(0.10f, 0.80f, n.y);
        d = F4{scl(d.g, up), d.w * up};
    }
    if (f.pattern == 12u) d.w *= smoothstepf(0.45f, 0.95f, n.y);
    const float value = d.w;
    const float amount = f.strength * value;
    albedo = mix3(albedo, ld3(f.tint), amount);
    if (f.pattern == 9u) {
        rough += fmaxf(f.roughness_delta, 0.0f) * amount;
        metallic *= 1.0f - amount;
        spec = mixf(spec, 0.025f, amount);
    } else if (f.pattern == 12u) {
        rough = mixf(rough, clampT(rough + f.roughness_delta, 0.025f, 1.0f), amount);
        spec = mixf(spec, 0.0204f, amount);
        metallic *= 1.0f - amount;
    } else {
        rough += f.roughness_delta * (2.0f * value - 1.0f);
    }
    if (f.pattern == 2u) metallic *= 1.0f - 0.75f * amount;
    if (fabsf(f.bump) > 0.0f) {  // finish_normal: the slope's tangential part tilts the normal, at most 0.75
        const F3 slope = sub(d.g, scl(n, dot3(d.g, n)));
        const float limit = 0.75f / fmaxf(len3(slope), 1e-5f);
        shading_n = safe_normalize(sub(n, scl(slope, clampT(f.bump, -limit, limit))));
    }
}

}  // namespace syf
