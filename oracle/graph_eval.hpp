// Oracle: CPU restatement of the reference's node-graph interpreter (TEST INFRASTRUCTURE ONLY,
// included by syf_oracle.cpp inside its anonymous namespace after the vector helpers, the
// spherical and trigonometric powers and power_node).
//
// It follows crates/syfrack-core/src/scene/graph/eval.rs literally, over the graph as the
// syf_graph C structs describe it: a plan of node ids in topological order (validate.rs:45-84,
// eval.rs:33-92), a map from node id to its Point or Distance value, the point operations
// applied through apply_seeded with the Iterate recursion (eval.rs:346-556), the fences
// (:558-599), the combinations (:601-657), evaluate_raw (:659-826) and the normalisation of
// PreparedGraph::evaluate (:98-114); iteration_product (validate.rs:112-137), trap_scales
// (eval.rs:853-877) and the Lipschitz factor and step scale (eval.rs:828-849, 879-881).  It does
// not validate beyond what evaluation needs (ids resolve, no cycle): the library's validation
// is checked against the reference's own rules in the tests.  Templated on the arithmetic type:
// float is the reference's f32, double the FP64 twin.
//
// glam is not available here.  Rotate builds the quaternion Rx * Ry * Rz from half-angle
// rotations and rotates with glam's Quat * Vec3 form (v (w^2 - |b|^2) + b 2 (v.b) + (b x v) 2w):
// the convention is pinned by the reference's own test (eval.rs:958-969), the rounding of glam's
// Euler conversion is not (a few ulps).

struct ONode {
    uint32_t id = 0, kind = 0, flags = 0;
    bool enabled = true;
    float v[6] = {0, 0, 0, 0, 0, 0};
    uint32_t u[3] = {0, 0, 0};
    std::vector<uint32_t> body;
};

struct OGraph {
    std::vector<ONode> nodes;
    std::vector<syf_graph_link> links;
    float safety = 1.0f, hint = 2.0f;
    // prepared (PreparedGraph::new)
    uint32_t output = 0;
    std::vector<uint32_t> plan;
    float trap_scales[4] = {1, 1, 1, 1};
    uint32_t iteration_product = 1;
    bool ok = false;

    const ONode* node(uint32_t id) const {
        for (const ONode& n : nodes)
            if (n.id == id) return &n;
        return nullptr;
    }
    // Graph::input_of: the first link into (to, port)
    bool input_of(uint32_t to, uint32_t port, uint32_t* from) const {
        for (const syf_graph_link& l : links)
            if (l.to_node == to && l.to_port == port) {
                *from = l.from_node;
                return true;
            }
        return false;
    }
    const std::vector<uint32_t>* body_of(uint32_t id) const {
        const ONode* n = node(id);
        return (n && n->kind == SYF_NODE_ITERATE) ? &n->body : nullptr;
    }
};

inline bool is_point_to_point(uint32_t k) { return k >= SYF_NODE_TRANSLATE && k <= SYF_NODE_TRAP; }
inline bool is_fence(uint32_t k) { return k >= SYF_NODE_SPHERE && k <= SYF_NODE_BULB_FENCE; }
inline bool is_combine(uint32_t k) { return k >= SYF_NODE_UNION && k <= SYF_NODE_SMOOTH_INTERSECTION; }
// NodeKind::inputs().len() (mod.rs:348-356)
inline uint32_t input_count(uint32_t k) {
    if (k == SYF_NODE_INPUT) return 0;
    if (k == SYF_NODE_BOUNDED || is_combine(k)) return 2;
    return 1;
}

// active_nodes (validate.rs:24-41): the nodes that reach the Output along links, sorted.
inline std::vector<uint32_t> active_nodes(const OGraph& g) {
    std::vector<uint32_t> seen;
    const ONode* out = nullptr;
    for (const ONode& n : g.nodes)
        if (n.kind == SYF_NODE_OUTPUT) {
            out = &n;
            break;
        }
    if (!out) return seen;
    std::vector<uint32_t> stack{out->id};
    while (!stack.empty()) {
        uint32_t id = stack.back();
        stack.pop_back();
        if (std::find(seen.begin(), seen.end(), id) != seen.end()) continue;
        seen.push_back(id);
        for (const syf_graph_link& l : g.links)
            if (l.to_node == id) stack.push_back(l.from_node);
    }
    std::sort(seen.begin(), seen.end());
    return seen;
}

inline bool is_body_node(const OGraph& g, uint32_t id) {
    for (const ONode& n : g.nodes)
        if (n.kind == SYF_NODE_ITERATE && std::find(n.body.begin(), n.body.end(), id) != n.body.end()) return true;
    return false;
}

// topo_order (validate.rs:45-84): Kahn's algorithm over the active non-body nodes, the
// smallest ready id first; false on a cycle.
inline bool topo_order(const OGraph& g, std::vector<uint32_t>* order) {
    std::vector<uint32_t> act = active_nodes(g);
    std::vector<std::pair<uint32_t, size_t>> indeg;  // (id, in-degree)
    for (uint32_t id : act)
        if (!is_body_node(g, id)) indeg.push_back({id, 0});
    auto find = [&](uint32_t id) -> std::pair<uint32_t, size_t>* {
        for (auto& e : indeg)
            if (e.first == id) return &e;
        return nullptr;
    };
    for (const syf_graph_link& l : g.links)
        if (find(l.to_node) && std::find(act.begin(), act.end(), l.from_node) != act.end()) find(l.to_node)->second++;
    std::vector<uint32_t> ready;
    for (auto& e : indeg)
        if (e.second == 0) ready.push_back(e.first);
    auto desc = [](uint32_t a, uint32_t b) { return a > b; };
    std::sort(ready.begin(), ready.end(), desc);
    order->clear();
    while (!ready.empty()) {
        uint32_t id = ready.back();
        ready.pop_back();
        order->push_back(id);
        std::vector<uint32_t> next;
        for (const syf_graph_link& l : g.links)
            if (l.from_node == id) {
                auto* e = find(l.to_node);
                if (e && --e->second == 0) next.push_back(l.to_node);
            }
        ready.insert(ready.end(), next.begin(), next.end());
        std::sort(ready.begin(), ready.end(), desc);
    }
    return order->size() == indeg.size();
}

// iteration_product (validate.rs:112-137)
inline uint32_t product_of(const OGraph& g, uint32_t id, std::vector<uint32_t>& seen) {
    if (std::find(seen.begin(), seen.end(), id) != seen.end()) return UINT32_MAX;
    seen.push_back(id);
    uint32_t p = 1;
    const ONode* n = g.node(id);
    if (n && n->kind == SYF_NODE_ITERATE) {
        uint32_t m = 1;
        bool any = false;
        for (uint32_t b : n->body) {
            uint32_t q = product_of(g, b, seen);
            m = any ? std::max(m, q) : q;
            any = true;
        }
        uint64_t prod = uint64_t(n->u[0]) * uint64_t(m);
        p = prod > UINT32_MAX ? UINT32_MAX : uint32_t(prod);
    }
    seen.pop_back();
    return p;
}
inline uint32_t iteration_product(const OGraph& g) {
    uint32_t best = 1;
    bool any = false;
    for (uint32_t id : active_nodes(g)) {
        if (is_body_node(g, id)) continue;
        std::vector<uint32_t> seen;
        uint32_t p = product_of(g, id, seen);
        best = any ? std::max(best, p) : p;
        any = true;
    }
    return best;
}

// trap_scales (eval.rs:853-877)
inline void trap_walk(const OGraph& g, uint32_t id, uint32_t depth, float* scales) {
    if (depth > 3u + 1u) return;
    const ONode* n = g.node(id);
    if (!n || !n->enabled) return;
    if (n->kind == SYF_NODE_TRAP) scales[std::min(n->u[0], 3u)] = n->v[0];
    if (n->kind == SYF_NODE_ITERATE)
        for (uint32_t b : n->body) trap_walk(g, b, depth + 1, scales);
}

// lipschitz (eval.rs:828-849): the product of sqrt(1 + (rate * radius)^2) over the enabled
// Twist and Bend nodes of the active nodes and their bodies.
inline void lip_visit(const OGraph& g, uint32_t id, float radius, uint32_t depth, float* l) {
    if (depth > 3u + 1u) return;
    const ONode* n = g.node(id);
    if (!n || !n->enabled) return;
    if (n->kind == SYF_NODE_TWIST || n->kind == SYF_NODE_BEND) {
        float x = n->v[0] * radius;
        *l = *l * std::sqrt(1.0f + x * x);
    }
    if (const std::vector<uint32_t>* body = g.body_of(id))
        for (uint32_t b : *body) lip_visit(g, b, radius, depth + 1, l);
}
inline float lipschitz(const OGraph& g, float radius) {
    float l = 1.0f;
    for (uint32_t id : active_nodes(g)) lip_visit(g, id, radius, 0, &l);
    return l;
}
inline float step_scale(const OGraph& g) { return g.safety / lipschitz(g, g.hint); }

// PreparedGraph::new (eval.rs:33-92) without instances: the plan of the Output.
inline bool prepare(OGraph& g) {
    std::vector<uint32_t> order;
    g.ok = false;
    if (!topo_order(g, &order)) return false;
    g.output = 0;
    for (const ONode& n : g.nodes)
        if (n.kind == SYF_NODE_OUTPUT) {
            g.output = n.id;
            break;
        }
    std::vector<uint32_t> needed;
    std::vector<uint32_t> stack{g.output};
    while (!stack.empty()) {
        uint32_t id = stack.back();
        stack.pop_back();
        if (std::find(needed.begin(), needed.end(), id) != needed.end()) continue;
        needed.push_back(id);
        const ONode* n = g.node(id);
        if (!n) continue;
        for (uint32_t port = 0; port < input_count(n->kind); ++port) {
            uint32_t src;
            if (g.input_of(id, port, &src)) stack.push_back(src);
        }
    }
    g.plan.clear();
    for (uint32_t id : order)
        if (std::find(needed.begin(), needed.end(), id) != needed.end()) g.plan.push_back(id);
    for (float& s : g.trap_scales) s = 1.0f;
    for (uint32_t id : order) trap_walk(g, id, 0, g.trap_scales);
    g.iteration_product = iteration_product(g);
    g.ok = true;
    return true;
}

// ---------------------------------------------------------------- evaluation

template <class T> struct PState {
    V3<T> p;
    T dr;
    uint32_t iter;
    T traps[4];
};
template <class T> struct DState {
    T d;
    uint32_t material, iter;
    T traps[4];
};

template <class T> inline T pi_over_180();
template <> inline float pi_over_180<float>() { return 3.14159265358979323846f / 180.0f; }  // f32::to_radians
template <> inline double pi_over_180<double>() { return 3.14159265358979323846 / 180.0; }
// f32::to_radians (deg * (PI / 180) in f32); the FP64 twin converts in double
template <class T> inline T to_radians(float deg) { return T(deg) * pi_over_180<T>(); }
template <class T> inline T tau();
template <> inline float tau<float>() { return 6.28318530717958647692f; }
template <> inline double tau<double>() { return 6.28318530717958647692; }

// Rust rem_euclid for floats: r = a % b; r < 0 ? r + |b| : r
template <class T> inline T rem_euclid(T a, T b) {
    T r = std::fmod(a, b);
    return r < T(0) ? r + std::fabs(b) : r;
}
// Vec3::try_normalize().unwrap_or(Y), in f32 as the reference
inline V3<float> try_normalize_or_y(const float* v) {
    V3<float> n{v[0], v[1], v[2]};
    float rcp = 1.0f / length(n);
    if (std::isfinite(rcp) && rcp > 0.0f) return n * rcp;
    return V3<float>{0.0f, 1.0f, 0.0f};
}
template <class T> inline V3<T> vcast(V3<float> a) { return v3(T(a.x), T(a.y), T(a.z)); }
template <class T> inline V3<T> vparam(const float* v) { return v3(T(v[0]), T(v[1]), T(v[2])); }
template <class T> inline T max_element(V3<T> a) { return fmax_(a.x, fmax_(a.y, a.z)); }
template <class T> inline T min_element(V3<T> a) { return fmin_(a.x, fmin_(a.y, a.z)); }
template <class T> inline T len2(T x, T y) { return std::sqrt(x * x + y * y); }

// rot2 (eval.rs:122-125)
template <class T> inline Vec2<T> rot2(T x, T y, T a) {
    T s = std::sin(a), c = std::cos(a);
    return Vec2<T>{c * x - s * y, s * x + c * y};
}

// Quat::from_euler(XYZ, a, b, c) as Rx(a) Ry(b) Rz(c) and Quat * Vec3.
template <class T> struct Quat {
    T x, y, z, w;
};
template <class T> inline Quat<T> qmul(Quat<T> a, Quat<T> b) {
    return Quat<T>{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                   a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class T> Quat<T> euler_xyz(T a, T b, T c) {
    T ha = a * T(0.5), hb = b * T(0.5), hc = c * T(0.5);
    Quat<T> qx{std::sin(ha), T(0), T(0), std::cos(ha)};
    Quat<T> qy{T(0), std::sin(hb), T(0), std::cos(hb)};
    Quat<T> qz{T(0), T(0), std::sin(hc), std::cos(hc)};
    return qmul(qmul(qx, qy), qz);
}
template <class T> inline V3<T> qrotate(Quat<T> q, V3<T> v) {
    V3<T> b = v3(q.x, q.y, q.z);
    T b2 = dot(b, b);
    return (v * (q.w * q.w - b2) + b * (dot(v, b) * T(2))) + cross(b, v) * (q.w * T(2));
}

// terrain_hash / terrain_noise / rock_distance (eval.rs:235-273)
inline uint32_t terrain_hash_u(int32_t cx, int32_t cy, int32_t cz, uint32_t seed) {
    uint32_t h = (uint32_t(cx) * 0x8da6b343u) ^ (uint32_t(cy) * 0xd8163841u) ^ (uint32_t(cz) * 0xcb1ab31fu) ^ seed;
    h = (h ^ (h >> 16)) * 0x7feb352du;
    h = (h ^ (h >> 15)) * 0x846ca68bu;
    h ^= h >> 16;
    return h;
}
// Rust `as i32` from a float: saturating, NaN -> 0
template <class T> inline int32_t sat_i32(T v) {
    if (v != v) return 0;
    if (v <= T(-2147483648.0)) return INT32_MIN;
    if (v >= T(2147483647.0)) return INT32_MAX;
    return int32_t(v);
}
template <class T> inline T terrain_hash(int32_t x, int32_t y, int32_t z, uint32_t seed) {
    return T(terrain_hash_u(x, y, z, seed) >> 8) / T(16777216.0f);
}
template <class T> inline T terrain_noise(V3<T> p, uint32_t seed) {
    V3<T> base = v3(std::floor(p.x), std::floor(p.y), std::floor(p.z));
    V3<T> f = p - base;
    auto sm = [](T x) { return x * x * (T(3) - T(2) * x); };
    V3<T> t = v3(sm(f.x), sm(f.y), sm(f.z));
    int32_t c0 = sat_i32(base.x), c1 = sat_i32(base.y), c2 = sat_i32(base.z);
    auto mix = [](T a, T b, T w) { return a + (b - a) * w; };
    auto h = [&](int32_t x, int32_t y, int32_t z) {
        return terrain_hash<T>(int32_t(uint32_t(c0) + uint32_t(x)), int32_t(uint32_t(c1) + uint32_t(y)),
                               int32_t(uint32_t(c2) + uint32_t(z)), seed);
    };
    T a = mix(h(0, 0, 0), h(1, 0, 0), t.x);
    T b = mix(h(0, 1, 0), h(1, 1, 0), t.x);
    T d = mix(h(0, 0, 1), h(1, 0, 1), t.x);
    T e = mix(h(0, 1, 1), h(1, 1, 1), t.x);
    return mix(mix(a, b, t.y), mix(d, e, t.y), t.z);
}
template <class T> inline T rock_distance(V3<T> p, V3<T> half, T erosion, uint32_t seed) {
    T h = min_element(half);
    V3<T> q = p * (T(1.15f) / h);
    T relief = T(0.57f) * terrain_noise(q, seed) + T(0.29f) * terrain_noise(q * T(2.03f), seed + 97u) +
               T(0.14f) * terrain_noise(q * T(4.17f), seed + 193u) - T(0.5f);
    V3<T> ph = v3(p.x / half.x, p.y / half.y, p.z / half.z);
    return (length(ph) - T(1) - erosion * relief) * h / (T(1) + T(5.22f) * erosion);
}

// measure (eval.rs:309-331)
template <class T> inline T measure(uint32_t m, V3<T> p) {
    switch (m) {
    case SYF_MEASURE_NORM: return length(p);
    case SYF_MEASURE_AXIS_X: return len2(p.y, p.z);
    case SYF_MEASURE_AXIS_Y: return len2(p.x, p.z);
    case SYF_MEASURE_AXIS_Z: return len2(p.x, p.y);
    case SYF_MEASURE_PLANE_Y: return std::fabs(p.y);
    case SYF_MEASURE_ANGLE_XY: return (p.x == T(0) && p.y == T(0)) ? T(0.5) : std::atan2(p.y, p.x) / tau<T>() + T(0.5);
    case SYF_MEASURE_ANGLE_XZ: return (p.x == T(0) && p.z == T(0)) ? T(0.5) : std::atan2(p.z, p.x) / tau<T>() + T(0.5);
    case SYF_MEASURE_ANGLE_YZ: return (p.y == T(0) && p.z == T(0)) ? T(0.5) : std::atan2(p.z, p.y) / tau<T>() + T(0.5);
    default: return std::log1p(length(p));  // LogNorm
    }
}

// apply_seeded (eval.rs:346-556); c is the Input point, seed the (point, dr) that AddSeed adds.
template <class T>
PState<T> apply_seeded(const OGraph& g, uint32_t id, PState<T> s, V3<T> c, V3<T> seed_p, T seed_dr) {
    const ONode* n = g.node(id);
    if (!n || !n->enabled) return s;
    const float* v = n->v;
    switch (n->kind) {
    case SYF_NODE_TRANSLATE: s.p = s.p + vparam<T>(v); break;
    case SYF_NODE_ROTATE: {
        // the quaternion in f32, as the reference builds it (the FP64 twin rotates by it in double)
        const Quat<float> qf = euler_xyz(to_radians<float>(v[0]), to_radians<float>(v[1]), to_radians<float>(v[2]));
        s.p = qrotate(Quat<T>{T(qf.x), T(qf.y), T(qf.z), T(qf.w)}, s.p);
        break;
    }
    case SYF_NODE_SCALE:
        s.p = s.p * T(v[0]);
        s.dr = s.dr * std::fabs(T(v[0]));
        break;
    case SYF_NODE_SCALE_AXES: {
        V3<T> f = vparam<T>(v);
        s.p = v3(s.p.x * f.x, s.p.y * f.y, s.p.z * f.z);
        s.dr = s.dr * max_element(f);
        break;
    }
    case SYF_NODE_MIRROR: {
        V3<T> nn = vcast<T>(try_normalize_or_y(v));
        T h = dot(s.p, nn) - T(v[3]);
        if (h < T(0)) s.p = s.p - nn * (T(2) * h);
        break;
    }
    case SYF_NODE_BOX_FOLD: {
        T l = T(v[0]);
        s.p = v3(clampv(s.p.x, -l, l) * T(2) - s.p.x, clampv(s.p.y, -l, l) * T(2) - s.p.y,
                 clampv(s.p.z, -l, l) * T(2) - s.p.z);
        break;
    }
    case SYF_NODE_BOX_FOLD_AXES: {
        V3<T> l = vparam<T>(v);
        V3<T> f = v3(clampv(s.p.x, -l.x, l.x) * T(2) - s.p.x, clampv(s.p.y, -l.y, l.y) * T(2) - s.p.y,
                     clampv(s.p.z, -l.z, l.z) * T(2) - s.p.z);
        if (n->flags & 1u) s.p.x = f.x;
        if (n->flags & 2u) s.p.y = f.y;
        if (n->flags & 4u) s.p.z = f.z;
        break;
    }
    case SYF_NODE_SPHERE_INVERSION: {
        T r = T(v[0]), m = T(v[1]);
        T k = r * r / fmax_(dot(s.p, s.p), m * m);
        s.p = s.p * k;
        s.dr = s.dr * k;
        break;
    }
    case SYF_NODE_SPHERE_FOLD: {
        T r2 = dot(s.p, s.p);
        T mr2 = T(v[0]) * T(v[0]);
        T fr2 = T(v[1]) * T(v[1]);
        T k = r2 < mr2 ? fr2 / mr2 : (r2 < fr2 ? fr2 / r2 : T(1));
        s.p = s.p * k;
        s.dr = s.dr * k;
        break;
    }
    case SYF_NODE_ABS:
        if (n->flags & 1u) s.p.x = std::fabs(s.p.x);
        if (n->flags & 2u) s.p.y = std::fabs(s.p.y);
        if (n->flags & 4u) s.p.z = std::fabs(s.p.z);
        break;
    case SYF_NODE_MENGER_FOLD: {
        V3<T> z = vabs(s.p);
        if (z.x < z.y) z = v3(z.y, z.x, z.z);
        if (z.x < z.z) z = v3(z.z, z.y, z.x);
        if (z.y < z.z) z = v3(z.x, z.z, z.y);
        T sc = T(v[0]);
        V3<T> off = v3(T(v[1]), T(v[2]), T(v[3]));
        z = z * sc - off * (sc - T(1));
        if (z.z < T(-0.5) * off.z * (sc - T(1))) z.z = z.z + off.z * (sc - T(1));
        s.p = z;
        s.dr = s.dr * sc;
        break;
    }
    case SYF_NODE_KALEIDO: {
        // sector and phase in f32, as the reference computes them
        T sector = T(tau<float>() / float(std::max(n->u[0], 1u)));
        T a = std::atan2(s.p.z, s.p.x) + T(to_radians<float>(v[0]));
        a = std::fabs(rem_euclid(a + sector * T(0.5), sector) - sector * T(0.5));
        T r = len2(s.p.x, s.p.z);
        s.p.x = r * std::cos(a);
        s.p.z = r * std::sin(a);
        break;
    }
    case SYF_NODE_POWER: power_node(&s.p, &s.dr, T(v[0])); break;
    case SYF_NODE_ADD_ORIGIN:
        s.p = s.p + c;
        s.dr = s.dr + T(1);
        break;
    case SYF_NODE_ADD_SEED: {
        V3<T> m = vparam<T>(v);
        s.p = s.p + v3(seed_p.x * m.x, seed_p.y * m.y, seed_p.z * m.z);
        s.dr = s.dr + seed_dr * max_element(vabs(m));
        break;
    }
    case SYF_NODE_ADD_CONSTANT: {
        V3<T> m = vparam<T>(v), j = vparam<T>(v + 3);
        s.p = s.p + v3(c.x * m.x + j.x, c.y * m.y + j.y, c.z * m.z + j.z);
        s.dr = s.dr + max_element(vabs(m));
        break;
    }
    case SYF_NODE_TWIST: {
        Vec2<T> r = rot2(s.p.x, s.p.z, T(v[0]) * s.p.y);
        s.p.x = r.x;
        s.p.z = r.y;
        break;
    }
    case SYF_NODE_BEND: {
        Vec2<T> r = rot2(s.p.x, s.p.y, T(v[0]) * s.p.x);
        s.p.x = r.x;
        s.p.y = r.y;
        break;
    }
    case SYF_NODE_TILE: {
        V3<T> per = vparam<T>(v);
        V3<T> half = per * T(0.5);
        s.p = v3(rem_euclid(s.p.x + half.x, per.x) - half.x, rem_euclid(s.p.y + half.y, per.y) - half.y,
                 rem_euclid(s.p.z + half.z, per.z) - half.z);
        break;
    }
    case SYF_NODE_REPEAT_FINITE: {
        V3<T> per = vparam<T>(v);
        V3<T> last = v3(T(float(n->u[0])), T(float(n->u[1])), T(float(n->u[2]))) - v3(T(1), T(1), T(1));
        V3<T> half = last * T(0.5);
        auto cell = [](T p, T per_, T half_, T last_) {
            return clampv(std::floor(p / per_ + half_ + T(0.5)), T(0), last_) - half_;
        };
        V3<T> cl = v3(cell(s.p.x, per.x, half.x, last.x), cell(s.p.y, per.y, half.y, last.y),
                      cell(s.p.z, per.z, half.z, last.z));
        s.p = s.p - v3(cl.x * per.x, cl.y * per.y, cl.z * per.z);
        break;
    }
    case SYF_NODE_ITERATE: {
        V3<T> lp = s.p;
        T ldr = s.dr;
        for (uint32_t k = 0; k < n->u[0]; ++k) {
            if (length(s.p) > T(v[0])) break;
            for (uint32_t b : n->body) s = apply_seeded(g, b, s, c, lp, ldr);
            s.iter += 1;
        }
        break;
    }
    case SYF_NODE_TRAP: {
        uint32_t i = std::min(n->u[0], 3u);
        T m = measure(n->u[1], s.p);
        if (n->u[2] == SYF_TRAP_MINIMUM) s.traps[i] = fmin_(s.traps[i], m);
        else if (n->u[2] == SYF_TRAP_MAXIMUM) s.traps[i] = s.traps[i] >= T(1e30) ? m : fmax_(s.traps[i], m);
        else s.traps[i] = m;
        break;
    }
    default: break;
    }
    return s;
}

// fence (eval.rs:558-599)
template <class T> T fence(const ONode& n, const PState<T>& s) {
    const float* v = n.v;
    V3<T> p = s.p;
    switch (n.kind) {
    case SYF_NODE_SPHERE: return (length(p) - T(v[0])) / s.dr;
    case SYF_NODE_TORUS: return (len2(len2(p.x, p.z) - T(v[0]), p.y) - T(v[1])) / s.dr;
    case SYF_NODE_CYLINDER: {
        T qx = len2(p.x, p.z) - T(v[0]), qy = std::fabs(p.y) - T(v[1]);
        return (len2(fmax_(qx, T(0)), fmax_(qy, T(0))) + fmin_(fmax_(qx, qy), T(0))) / s.dr;
    }
    case SYF_NODE_BOX:
    case SYF_NODE_BOUNDED: {
        V3<T> q = vabs(p) - vparam<T>(v);
        V3<T> qm = v3(fmax_(q.x, T(0)), fmax_(q.y, T(0)), fmax_(q.z, T(0)));
        return (length(qm) + fmin_(fmax_(q.x, fmax_(q.y, q.z)), T(0))) / s.dr;
    }
    case SYF_NODE_ROUND_BOX: {
        V3<T> half = vparam<T>(v);
        T r = fmin_(T(v[3]), min_element(half));
        V3<T> q = vabs(p) - (half - v3(r, r, r));
        V3<T> qm = v3(fmax_(q.x, T(0)), fmax_(q.y, T(0)), fmax_(q.z, T(0)));
        return (length(qm) + fmin_(max_element(q), T(0)) - r) / s.dr;
    }
    case SYF_NODE_ROCK: return rock_distance(p, vparam<T>(v), T(v[3]), n.u[0]) / s.dr;
    case SYF_NODE_PLANE: return (dot(p, vcast<T>(try_normalize_or_y(v))) - T(v[3])) / s.dr;
    case SYF_NODE_NORM_FENCE: return (length(p) - T(v[0])) / std::fabs(s.dr);
    case SYF_NODE_BULB_FENCE: {
        T r = fmax_(length(p), T(1e-12));
        return T(0.5) * std::log(r) * r / s.dr;
    }
    default: return std::numeric_limits<T>::infinity();
    }
}

// combine (eval.rs:601-657)
template <class T> DState<T> combine(const ONode& n, const DState<T>& a, const DState<T>& b) {
    auto empty = [](T d) { return d >= T(1e30); };
    auto blend = [&](T h, T d) {
        DState<T> r;
        for (int i = 0; i < 4; ++i) {
            bool fa = std::isfinite(a.traps[i]), fb = std::isfinite(b.traps[i]);
            r.traps[i] = (fa && fb) ? b.traps[i] + (a.traps[i] - b.traps[i]) * h
                                    : fa ? a.traps[i] : fb ? b.traps[i] : std::numeric_limits<T>::infinity();
        }
        r.d = d;
        r.material = h >= T(0.5) ? a.material : b.material;
        r.iter = h >= T(0.5) ? a.iter : b.iter;
        return r;
    };
    const T k = T(n.v[0]);
    switch (n.kind) {
    case SYF_NODE_UNION: {
        DState<T> r = a.d <= b.d ? a : b;
        r.d = fmin_(a.d, b.d);
        return r;
    }
    case SYF_NODE_INTERSECTION: {
        DState<T> r = a.d >= b.d ? a : b;
        r.d = fmax_(a.d, b.d);
        return r;
    }
    case SYF_NODE_SUBTRACTION: {
        DState<T> r = a;
        r.d = fmax_(a.d, -b.d);
        return r;
    }
    case SYF_NODE_SMOOTH_UNION: {
        if (empty(a.d)) return b;
        if (empty(b.d)) return a;
        T h = std::min(std::max(T(0.5) + T(0.5) * (b.d - a.d) / k, T(0)), T(1));
        return blend(h, b.d + (a.d - b.d) * h - k * h * (T(1) - h));
    }
    case SYF_NODE_SMOOTH_INTERSECTION: {
        if (empty(a.d)) {
            DState<T> r = b;
            r.d = std::numeric_limits<T>::infinity();
            return r;
        }
        if (empty(b.d)) {
            DState<T> r = a;
            r.d = std::numeric_limits<T>::infinity();
            return r;
        }
        T h = std::min(std::max(T(0.5) - T(0.5) * (b.d - a.d) / k, T(0)), T(1));
        return blend(h, b.d + (a.d - b.d) * h + k * h * (T(1) - h));
    }
    default: return a;
    }
}

// evaluate_raw (eval.rs:659-826) for the Output's plan, then the normalisation of
// PreparedGraph::evaluate (eval.rs:98-114).
template <class T> DeOut<T> graph_evaluate(const OGraph& g, V3<T> p) {
    const T INF = std::numeric_limits<T>::infinity();
    DState<T> empty{INF, 0u, 0u, {INF, INF, INF, INF}};
    PState<T> start{p, T(1), 0u, {INF, INF, INF, INF}};
    struct Val {
        uint32_t id;
        bool is_point;
        PState<T> ps;
        DState<T> ds;
    };
    std::vector<Val> values;
    values.reserve(g.plan.size());
    auto lookup = [&](uint32_t id) -> const Val* {
        for (const Val& x : values)
            if (x.id == id) return &x;
        return nullptr;
    };
    uint32_t runs = 0;
    for (uint32_t id : g.plan) {
        const ONode& n = *g.node(id);
        auto get_point = [&](uint32_t port) {
            uint32_t src;
            if (g.input_of(id, port, &src))
                if (const Val* x = lookup(src))
                    if (x->is_point) return x->ps;
            return start;
        };
        auto get_dist = [&](uint32_t port) {
            uint32_t src;
            if (g.input_of(id, port, &src))
                if (const Val* x = lookup(src))
                    if (!x->is_point) return x->ds;
            return empty;
        };
        Val out{id, false, start, empty};
        if (n.kind == SYF_NODE_INPUT) {
            out.is_point = true;
            out.ps = start;
        } else if (n.kind == SYF_NODE_OUTPUT) {
            out.ds = get_dist(0);
        } else if (is_point_to_point(n.kind)) {
            PState<T> s = get_point(0);
            out.is_point = true;
            out.ps = apply_seeded(g, id, s, p, s.p, s.dr);
            if (n.kind == SYF_NODE_ITERATE && n.enabled) runs += out.ps.iter - s.iter;
        } else if (is_fence(n.kind)) {
            PState<T> s = get_point(0);
            out.ds = DState<T>{n.enabled ? fence(n, s) : INF, 0u, s.iter, {s.traps[0], s.traps[1], s.traps[2], s.traps[3]}};
        } else if (n.kind == SYF_NODE_MATERIAL_TAG) {
            out.ds = get_dist(0);
            if (n.enabled) out.ds.material = n.u[0];
        } else if (n.kind == SYF_NODE_BOUNDED) {
            out.ds = get_dist(1);
            if (n.enabled) {
                PState<T> s = get_point(0);
                out.ds.d = fmax_(out.ds.d, fence(n, s));
            }
        } else if (!n.enabled) {
            out.ds = get_dist(0);  // a disabled Instances or combination passes its first input
        } else {
            out.ds = combine(n, get_dist(0), get_dist(1));
        }
        values.push_back(out);
    }
    DState<T> r = empty;
    if (const Val* x = lookup(g.output))
        if (!x->is_point) r = x->ds;
    DeOut<T> o;
    o.d = r.d;
    for (int i = 0; i < 4; ++i)
        o.traps[i] = std::isfinite(r.traps[i]) ? fmin_(r.traps[i] / T(g.trap_scales[i]), T(1)) : T(1);
    o.iter = std::min(r.iter, g.iteration_product);
    o.material = r.material;
    o.body_runs = runs;
    return o;
}

// Copies the C description into an OGraph and prepares it; false when ids do not resolve
// or the graph has a cycle.
inline bool load_graph(const syf_graph& cg, OGraph* g) {
    g->nodes.clear();
    g->links.assign(cg.links, cg.links + cg.link_count);
    g->safety = cg.safety;
    g->hint = cg.bounding_radius_hint;
    for (uint32_t i = 0; i < cg.node_count; ++i) {
        const syf_graph_node& s = cg.nodes[i];
        ONode n;
        n.id = s.id;
        n.kind = s.kind;
        n.flags = s.flags;
        n.enabled = s.enabled != 0;
        std::memcpy(n.v, s.v, sizeof n.v);
        std::memcpy(n.u, s.u, sizeof n.u);
        if (s.kind == SYF_NODE_ITERATE) {
            if (uint64_t(s.body_first) + s.body_count > cg.body_len) return false;
            n.body.assign(cg.bodies + s.body_first, cg.bodies + s.body_first + s.body_count);
        }
        g->nodes.push_back(n);
    }
    for (const syf_graph_link& l : g->links)
        if (!g->node(l.from_node) || !g->node(l.to_node)) return false;
    return prepare(*g);
}
