// Node graphs on the host: validation, derived properties and compilation into the register
// program that graph_device.hpp runs, and the C ABI entry points syf_graph_info and
// syf_job_create_graph.
//
// Validation restates Graph::validate_structure and Graph::validate (validate.rs:139-480) with
// the parameter bounds of Params for Graph (mod.rs:44-73, 527-662) and the English messages of
// the reference's catalogue (syfrack-i18n messages.tsv), prefixed "graph: " or "graph: node N: "
// like SceneError (scene/mod.rs:67-71).  The derived properties restate active_nodes and
// topo_order (validate.rs:24-84), iterate_depth and iteration_product (:86-137), trap_scales and
// the Lipschitz step scale (eval.rs:828-881).
//
// Compilation walks the values the Output needs (the plan of PreparedGraph::new, eval.rs:33-92)
// depth first from the Output.  Nodes that pass a value through (a disabled node, Output) are
// resolved to their source; a point value stays in the interpreter's current point until a
// node consumes it in place, and a value that still has consumers is saved to a slot first; a
// distance value likewise.  Evaluation order does not change a result: every node is a pure
// function of the evaluated point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>

#include <dlfcn.h>
#include <hip/hiprtc.h>

#include <filesystem>
#include <mutex>
#include <thread>

#include "api_internal.hpp"
#include "device.hpp"
#include "kernels.hpp"

namespace syf {
namespace {

constexpr uint32_t MAX_NODES = 1536;       // scene/graph/mod.rs:16
constexpr uint32_t MAX_DEPTH = 3;          // mod.rs:42
constexpr uint32_t MAX_ITER_PRODUCT = 256; // mod.rs:43

// English labels of the node kinds (NodeKind::label, mod.rs:263-308, through the catalogue).
const char* const KIND_LABEL[SYF_NODE_KINDS] = {
    "Input", "Output", "Translation", "Rotation", "Scale", "Axis Scale", "Mirror", "Box Fold", "Axis Fold",
    "Sphere Inversion", "Sphere Fold", "Absolute Value", "Menger Fold", "Kaleidoscope", "Polar Power", "Add Origin",
    "Reinject Local Point", "Fractal Constant", "Twist", "Bend", "Tile", "Finite Repeat", "Repeat", "Trap", "Sphere",
    "Torus", "Cylinder", "Box", "Rounded Box", "Eroded Rock", "Plane", "Norm Surface", "Bulb Surface",
    "Branch Material", "Detail Volume", "Independent Copies", "Union", "Intersection", "Subtraction", "Smooth Union",
    "Smooth Intersection"};

bool point_to_point(uint32_t k) { return k >= SYF_NODE_TRANSLATE && k <= SYF_NODE_TRAP; }
bool fence_kind(uint32_t k) { return k >= SYF_NODE_SPHERE && k <= SYF_NODE_BULB_FENCE; }
bool combine_kind(uint32_t k) { return k >= SYF_NODE_UNION && k <= SYF_NODE_SMOOTH_INTERSECTION; }

enum Port : uint8_t { PT_POINT, PT_DIST };
// NodeKind::inputs / outputs (mod.rs:348-367)
std::vector<Port> inputs_of(uint32_t k) {
    if (k == SYF_NODE_INPUT) return {};
    if (k == SYF_NODE_OUTPUT || k == SYF_NODE_MATERIAL_TAG || k == SYF_NODE_INSTANCES) return {PT_DIST};
    if (k == SYF_NODE_BOUNDED) return {PT_POINT, PT_DIST};
    if (combine_kind(k)) return {PT_DIST, PT_DIST};
    return {PT_POINT};
}
std::vector<Port> outputs_of(uint32_t k) {
    if (k == SYF_NODE_OUTPUT) return {};
    if (k == SYF_NODE_MATERIAL_TAG || k == SYF_NODE_BOUNDED || k == SYF_NODE_INSTANCES || fence_kind(k) ||
        combine_kind(k))
        return {PT_DIST};
    return {PT_POINT};
}
const char* port_name(Port p) { return p == PT_POINT ? "Point" : "Distance"; }

// Parameter bounds in the order Params for Graph visits them (mod.rs:527-662).  Forms: a float
// v[slot], a vector v[slot..slot+3] (.x .y .z), an integer u[slot], three integers u[0..3].
enum Form : uint8_t { F1, V3F, U1, U3 };
struct PSpec {
    const char* name;
    Form form;
    uint8_t slot;
    float lo, hi;
};
const float DIR = 1.0f;
std::vector<PSpec> params_of(uint32_t k) {
    switch (k) {
    case SYF_NODE_TRANSLATE: return {{"offset", V3F, 0, -10.0f, 10.0f}};
    case SYF_NODE_ROTATE: return {{"angles", V3F, 0, -180.0f, 180.0f}};
    case SYF_NODE_SCALE: return {{"factor", F1, 0, -4.0f, 4.0f}};
    case SYF_NODE_SCALE_AXES: return {{"factors", V3F, 0, 0.001f, 100.0f}};
    case SYF_NODE_MIRROR: return {{"normal", V3F, 0, -DIR, DIR}, {"offset", F1, 3, -10.0f, 10.0f}};
    case SYF_NODE_BOX_FOLD: return {{"limit", F1, 0, 0.5f, 2.0f}};
    case SYF_NODE_BOX_FOLD_AXES: return {{"limit", V3F, 0, 0.01f, 4.0f}};
    case SYF_NODE_SPHERE_INVERSION: return {{"radius", F1, 0, 0.01f, 10.0f}, {"min_radius", F1, 1, 0.0001f, 2.0f}};
    case SYF_NODE_ADD_CONSTANT: return {{"multiplier", V3F, 0, -4.0f, 4.0f}, {"julia", V3F, 3, -10.0f, 10.0f}};
    case SYF_NODE_ADD_SEED: return {{"multiplier", V3F, 0, -4.0f, 4.0f}};
    case SYF_NODE_SPHERE_FOLD: return {{"min_radius", F1, 0, 0.05f, 1.0f}, {"fixed_radius", F1, 1, 0.5f, 2.0f}};
    case SYF_NODE_MENGER_FOLD: return {{"scale", F1, 0, 2.0f, 4.0f}, {"offset", V3F, 1, 0.5f, 2.0f}};
    case SYF_NODE_KALEIDO: return {{"planes", U1, 0, 2.0f, 12.0f}, {"phase", F1, 0, -180.0f, 180.0f}};
    case SYF_NODE_POWER: return {{"power", F1, 0, 2.0f, 16.0f}};
    case SYF_NODE_TWIST:
    case SYF_NODE_BEND: return {{"rate", F1, 0, -2.0f, 2.0f}};
    case SYF_NODE_TILE: return {{"period", V3F, 0, 0.1f, 100.0f}};
    case SYF_NODE_INSTANCES: return {{"half", V3F, 0, 0.01f, 500.0f}, {"margin", F1, 3, 0.001f, 10.0f}};
    case SYF_NODE_REPEAT_FINITE: return {{"period", V3F, 0, 0.1f, 100.0f}, {"count", U3, 0, 1.0f, 32.0f}};
    case SYF_NODE_ITERATE: return {{"iterations", U1, 0, 1.0f, 64.0f}, {"bailout", F1, 0, 2.0f, 1.0e6f}};
    case SYF_NODE_TRAP: return {{"slot", U1, 0, 0.0f, 3.0f}, {"scale", F1, 0, 0.01f, 100.0f}};
    case SYF_NODE_SPHERE: return {{"radius", F1, 0, 0.01f, 100.0f}};
    case SYF_NODE_TORUS: return {{"radius", F1, 0, 0.01f, 100.0f}, {"tube", F1, 1, 0.01f, 100.0f}};
    case SYF_NODE_CYLINDER: return {{"radius", F1, 0, 0.01f, 100.0f}, {"half_height", F1, 1, 0.01f, 500.0f}};
    case SYF_NODE_MATERIAL_TAG: return {{"index", U1, 0, 0.0f, 8.0f}};
    case SYF_NODE_BOUNDED: return {{"half", V3F, 0, 0.01f, 500.0f}, {"margin", F1, 3, 0.001f, 10.0f}};
    case SYF_NODE_BOX: return {{"half", V3F, 0, 0.01f, 500.0f}};
    case SYF_NODE_ROUND_BOX: return {{"half", V3F, 0, 0.01f, 500.0f}, {"radius", F1, 3, 0.0f, 2.0f}};
    case SYF_NODE_ROCK:
        return {{"half", V3F, 0, 0.01f, 500.0f}, {"erosion", F1, 3, 0.0f, 0.6f}, {"seed", U1, 0, 0.0f, 65535.0f}};
    case SYF_NODE_PLANE: return {{"normal", V3F, 0, -DIR, DIR}, {"offset", F1, 3, -100.0f, 100.0f}};
    case SYF_NODE_NORM_FENCE: return {{"inset", F1, 0, 0.0f, 2.0f}};
    case SYF_NODE_SMOOTH_UNION:
    case SYF_NODE_SMOOTH_INTERSECTION: return {{"k", F1, 0, 0.001f, 10.0f}};
    default: return {};
    }
}

// Rust's Display of an f32: the shortest decimal that reads back to the same value, without an
// exponent.
std::string fmt_f32(float v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    for (int p = 1; p <= 9; ++p) {
        std::snprintf(buf, sizeof buf, "%.*g", p, double(v));
        if (std::strtof(buf, nullptr) == v) break;
    }
    if (!std::strchr(buf, 'e')) return buf;
    const double d = std::strtod(buf, nullptr);
    for (int decimals = 0; decimals <= 60; ++decimals) {
        std::snprintf(buf, sizeof buf, "%.*f", decimals, d);
        if (std::strtof(buf, nullptr) == v) break;
    }
    return buf;
}

struct Node {
    uint32_t id, kind, flags;
    bool enabled;
    float v[6];
    uint32_t u[3];
    std::vector<uint32_t> body;
};

// The checked graph, its plan and its derived properties.
struct Compiled {
    std::vector<GOp> prog;
    std::vector<uint32_t> node_of;  // the node of each instruction
    syf_fractal classic{};          // the classic family the graph is, when classic_family != SYF_FAMILY_GRAPH
    uint32_t classic_family = SYF_FAMILY_GRAPH;
    float step_scale = 1.0f;
    uint32_t iteration_product = 1;
    float trap_scales[4] = {1, 1, 1, 1};
    float hint = 2.0f;
};

struct Graph {
    std::vector<Node> nodes;
    std::vector<syf_graph_link> links;
    float safety = 1.0f, hint = 2.0f;
    std::map<uint32_t, size_t> index;

    const Node* node(uint32_t id) const {
        auto it = index.find(id);
        return it == index.end() ? nullptr : &nodes[it->second];
    }
    bool input_of(uint32_t to, uint32_t port, uint32_t* from) const {
        for (const syf_graph_link& l : links)
            if (l.to_node == to && l.to_port == port) {
                *from = l.from_node;
                return true;
            }
        return false;
    }
    const Node* output() const {
        for (const Node& n : nodes)
            if (n.kind == SYF_NODE_OUTPUT) return &n;
        return nullptr;
    }
    std::map<uint32_t, uint32_t> body_owners() const {
        std::map<uint32_t, uint32_t> owners;
        for (const Node& n : nodes)
            if (n.kind == SYF_NODE_ITERATE)
                for (uint32_t b : n.body) owners[b] = n.id;
        return owners;
    }
    // active_nodes (validate.rs:24-41)
    std::vector<uint32_t> active() const {
        std::vector<uint32_t> out;
        const Node* o = output();
        if (!o) return out;
        std::set<uint32_t> seen;
        std::vector<uint32_t> stack{o->id};
        while (!stack.empty()) {
            const uint32_t id = stack.back();
            stack.pop_back();
            if (!seen.insert(id).second) continue;
            for (const syf_graph_link& l : links)
                if (l.to_node == id) stack.push_back(l.from_node);
        }
        return std::vector<uint32_t>(seen.begin(), seen.end());
    }
    // topo_order (validate.rs:45-84): the smallest ready id first; false on a cycle
    bool topo_order(std::vector<uint32_t>* order) const {
        const std::vector<uint32_t> act = active();
        const std::set<uint32_t> act_set(act.begin(), act.end());
        const auto owners = body_owners();
        std::map<uint32_t, size_t> indeg;
        for (uint32_t id : act)
            if (!owners.count(id)) indeg[id] = 0;
        for (const syf_graph_link& l : links)
            if (indeg.count(l.to_node) && act_set.count(l.from_node)) indeg[l.to_node]++;
        std::set<uint32_t> ready;
        for (auto& e : indeg)
            if (e.second == 0) ready.insert(e.first);
        order->clear();
        while (!ready.empty()) {
            const uint32_t id = *ready.begin();
            ready.erase(ready.begin());
            order->push_back(id);
            for (const syf_graph_link& l : links)
                if (l.from_node == id) {
                    auto it = indeg.find(l.to_node);
                    if (it != indeg.end() && --it->second == 0) ready.insert(l.to_node);
                }
        }
        return order->size() == indeg.size();
    }
    // iterate_depth (validate.rs:86-109)
    uint32_t depth_of(uint32_t id, std::vector<uint32_t>& seen) const {
        if (std::find(seen.begin(), seen.end(), id) != seen.end()) return 99;
        seen.push_back(id);
        uint32_t d = 0;
        const Node* n = node(id);
        if (n && n->kind == SYF_NODE_ITERATE) {
            uint32_t m = 0;
            for (uint32_t b : n->body) m = std::max(m, depth_of(b, seen));
            d = 1 + m;
        }
        seen.pop_back();
        return d;
    }
    uint32_t iterate_depth() const {
        uint32_t d = 0;
        for (const Node& n : nodes) {
            std::vector<uint32_t> seen;
            d = std::max(d, depth_of(n.id, seen));
        }
        return d;
    }
    // iteration_product (validate.rs:112-137)
    uint32_t product_of(uint32_t id, std::vector<uint32_t>& seen) const {
        if (std::find(seen.begin(), seen.end(), id) != seen.end()) return UINT32_MAX;
        seen.push_back(id);
        uint32_t p = 1;
        const Node* n = node(id);
        if (n && n->kind == SYF_NODE_ITERATE) {
            uint32_t m = 1;
            bool any = false;
            for (uint32_t b : n->body) {
                const uint32_t q = product_of(b, seen);
                m = any ? std::max(m, q) : q;
                any = true;
            }
            const uint64_t prod = uint64_t(n->u[0]) * m;
            p = prod > UINT32_MAX ? UINT32_MAX : uint32_t(prod);
        }
        seen.pop_back();
        return p;
    }
    uint32_t iteration_product() const {
        const auto owners = body_owners();
        uint32_t best = 1;
        bool any = false;
        for (uint32_t id : active()) {
            if (owners.count(id)) continue;
            std::vector<uint32_t> seen;
            const uint32_t p = product_of(id, seen);
            best = any ? std::max(best, p) : p;
            any = true;
        }
        return best;
    }
    // trap_scales (eval.rs:853-877)
    void trap_walk(uint32_t id, uint32_t depth, float* scales) const {
        if (depth > MAX_DEPTH + 1) return;
        const Node* n = node(id);
        if (!n || !n->enabled) return;
        if (n->kind == SYF_NODE_TRAP) scales[std::min(n->u[0], 3u)] = n->v[0];
        if (n->kind == SYF_NODE_ITERATE)
            for (uint32_t b : n->body) trap_walk(b, depth + 1, scales);
    }
    // lipschitz (eval.rs:828-849)
    void lip_visit(uint32_t id, float radius, uint32_t depth, float* l) const {
        if (depth > MAX_DEPTH + 1) return;
        const Node* n = node(id);
        if (!n || !n->enabled) return;
        if (n->kind == SYF_NODE_TWIST || n->kind == SYF_NODE_BEND) {
            const float x = n->v[0] * radius;
            *l = *l * std::sqrt(1.0f + x * x);
        }
        if (n->kind == SYF_NODE_ITERATE)
            for (uint32_t b : n->body) lip_visit(b, radius, depth + 1, l);
    }
    float step_scale() const {
        float l = 1.0f;
        for (uint32_t id : active()) lip_visit(id, hint, 0, &l);
        return safety / l;
    }
};

template <class... A> int graph_err(const char* fmt, A... a) {
    char buf[768];
    std::snprintf(buf, sizeof buf, fmt, a...);
    return fail(SYF_E_INVALID, "graph: %s", buf);
}
template <class... A> int node_err(uint32_t id, const char* fmt, A... a) {
    char buf[768];
    std::snprintf(buf, sizeof buf, fmt, a...);
    return fail(SYF_E_INVALID, "graph: node %u: %s", id, buf);
}

int load(const syf_graph& g, Graph& G) {
    if (g.node_count && !g.nodes) return fail(SYF_E_INVALID, "graph: nodes must not be null");
    if (g.link_count && !g.links) return fail(SYF_E_INVALID, "graph: links must not be null");
    if (g.body_len && !g.bodies) return fail(SYF_E_INVALID, "graph: bodies must not be null");
    if (g.node_count > MAX_NODES) return graph_err("More than %u nodes", MAX_NODES);
    G.safety = g.safety;
    G.hint = g.bounding_radius_hint;
    G.links.assign(g.links, g.links + g.link_count);
    for (uint32_t i = 0; i < g.node_count; ++i) {
        const syf_graph_node& s = g.nodes[i];
        if (s.kind >= SYF_NODE_KINDS) return node_err(s.id, "unknown node kind %u", s.kind);
        Node n{s.id, s.kind, s.flags, s.enabled != 0, {}, {}, {}};
        std::memcpy(n.v, s.v, sizeof n.v);
        std::memcpy(n.u, s.u, sizeof n.u);
        if (s.kind == SYF_NODE_ITERATE) {
            if (uint64_t(s.body_first) + s.body_count > g.body_len)
                return node_err(s.id, "body_first + body_count (%u + %u) beyond the %u body entries", s.body_first,
                                s.body_count, g.body_len);
            n.body.assign(g.bodies + s.body_first, g.bodies + s.body_first + s.body_count);
        }
        if (s.kind == SYF_NODE_TRAP && s.u[1] > SYF_MEASURE_LOG_NORM) return node_err(s.id, "unknown trap measure %u", s.u[1]);
        if (s.kind == SYF_NODE_TRAP && s.u[2] > SYF_TRAP_LAST) return node_err(s.id, "unknown trap mode %u", s.u[2]);
        // Duplicate node identifier (validate.rs:150-158)
        if (!G.index.emplace(s.id, G.nodes.size()).second) return graph_err("Duplicate node identifier %u", s.id);
        G.nodes.push_back(std::move(n));
    }
    return SYF_OK;
}

// Kahn's algorithm over every node and link (validate.rs:412-438)
bool acyclic(const Graph& G) {
    std::map<uint32_t, size_t> indeg;
    for (const Node& n : G.nodes) indeg[n.id] = 0;
    for (const syf_graph_link& l : G.links)
        if (indeg.count(l.to_node)) indeg[l.to_node]++;
    std::vector<uint32_t> ready;
    for (auto& e : indeg)
        if (e.second == 0) ready.push_back(e.first);
    size_t done = 0;
    while (!ready.empty()) {
        const uint32_t id = ready.back();
        ready.pop_back();
        ++done;
        for (const syf_graph_link& l : G.links)
            if (l.from_node == id) {
                auto it = indeg.find(l.to_node);
                if (it != indeg.end() && --it->second == 0) ready.push_back(l.to_node);
            }
    }
    return done == indeg.size();
}

// Graph::validate_structure then Graph::validate (validate.rs:139-480).
int validate(const Graph& G) {
    size_t inputs = 0, outputs = 0;
    for (const Node& n : G.nodes) {
        inputs += n.kind == SYF_NODE_INPUT;
        outputs += n.kind == SYF_NODE_OUTPUT;
    }
    if (inputs != 1) return graph_err("The graph must have exactly one Input");
    if (outputs != 1) return graph_err("The graph must have exactly one Output");
    // parameters within their bounds, safety first (validate_params, params.rs:41-58)
    if (!(std::isfinite(G.safety) && G.safety >= 0.1f && G.safety <= 1.0f))
        return graph_err("parameter `fractal.safety` = %s is outside [0.1, 1]", fmt_f32(G.safety).c_str());
    for (const Node& n : G.nodes) {
        for (const PSpec& p : params_of(n.kind)) {
            const int count = (p.form == V3F || p.form == U3) ? 3 : 1;
            for (int k = 0; k < count; ++k) {
                const float val = p.form == F1 || p.form == V3F ? n.v[p.slot + k] : float(n.u[p.form == U3 ? k : p.slot]);
                if (std::isfinite(val) && val >= p.lo && val <= p.hi) continue;
                const char* axis = count == 3 ? (k == 0 ? ".x" : k == 1 ? ".y" : ".z") : "";
                return node_err(n.id, "parameter `fractal.nodes[%u].%s%s` = %s is outside [%s, %s]", n.id, p.name, axis,
                                fmt_f32(val).c_str(), fmt_f32(p.lo).c_str(), fmt_f32(p.hi).c_str());
            }
        }
    }
    for (const Node& n : G.nodes)
        if (n.kind == SYF_NODE_SCALE && std::fabs(n.v[0]) < 1e-3f)
            return node_err(n.id, "nodes[%u].factor: scale cannot be zero", n.id);
    // a node belongs to at most one body, once
    std::map<uint32_t, uint32_t> member_of;
    for (const Node& n : G.nodes) {
        if (n.kind != SYF_NODE_ITERATE) continue;
        for (uint32_t b : n.body) {
            auto ins = member_of.emplace(b, n.id);
            if (!ins.second) {
                const uint32_t first = ins.first->second;
                if (first == n.id) return node_err(b, "Node %u appears twice in the body of Repeat %u", b, first);
                return node_err(b, "Node %u is in two bodies (Repeat %u and Repeat %u)", b, first, n.id);
            }
        }
    }
    for (const Node& n : G.nodes) {
        if (n.kind != SYF_NODE_ITERATE) continue;
        for (uint32_t b : n.body) {
            const Node* m = G.node(b);
            if (!m) return node_err(n.id, "The body of Repeat %u references unknown node %u", n.id, b);
            if (!point_to_point(m->kind))
                return node_err(b, "The body of Repeat %u contains “%s”, which is not Point → Point", n.id,
                                KIND_LABEL[m->kind]);
            for (const syf_graph_link& l : G.links)
                if (l.from_node == b || l.to_node == b)
                    return node_err(b, "Node %u in the body of Repeat %u cannot have external links", b, n.id);
        }
    }
    for (const syf_graph_link& l : G.links) {
        const Node* from = G.node(l.from_node);
        if (!from) return graph_err("Link from unknown node %u", l.from_node);
        const Node* to = G.node(l.to_node);
        if (!to) return graph_err("Link to unknown node %u", l.to_node);
        const std::vector<Port> outs = outputs_of(from->kind), ins = inputs_of(to->kind);
        if (l.from_port >= outs.size()) return node_err(from->id, "Missing output port %u on node %u", l.from_port, from->id);
        if (l.to_port >= ins.size()) return node_err(to->id, "Missing input port %u on node %u", l.to_port, to->id);
        if (outs[l.from_port] != ins[l.to_port])
            return node_err(to->id, "Link of type %s to a %s port (%u → %u)", port_name(outs[l.from_port]),
                            port_name(ins[l.to_port]), from->id, to->id);
    }
    if (!acyclic(G)) return graph_err("Cycle in graph");
    std::set<std::pair<uint32_t, uint32_t>> seen_to;
    for (const syf_graph_link& l : G.links)
        if (!seen_to.insert({l.to_node, l.to_port}).second)
            return node_err(l.to_node, "Two links on port %u of node %u", l.to_port, l.to_node);
    const uint32_t depth = G.iterate_depth();
    if (depth > MAX_DEPTH) return graph_err("Repeat depth %u > %u", depth, MAX_DEPTH);
    const uint32_t product = G.iteration_product();
    if (product > MAX_ITER_PRODUCT) return graph_err("Iteration product %u > %u", product, MAX_ITER_PRODUCT);
    // validate(): every input port of every active node wired, the Output first
    const Node* out = G.output();
    uint32_t src;
    if (!G.input_of(out->id, 0, &src)) return node_err(out->id, "Output is not connected");
    std::vector<uint32_t> order;
    if (!G.topo_order(&order)) return graph_err("Cycle in graph");
    for (uint32_t id : order) {
        const Node* n = G.node(id);
        for (uint32_t port = 0; port < inputs_of(n->kind).size(); ++port)
            if (!G.input_of(id, port, &src))
                return node_err(id, "Input port %u of “node %u” is not connected", port, id);
    }
    return SYF_OK;
}

// ---------------------------------------------------------------- compilation

constexpr uint32_t VAL_START = 0xffffffffu;  // the Input point
constexpr uint32_t VAL_NONE = 0xfffffffeu;

struct Compiler {
    const Graph& G;
    std::vector<GOp> prog;
    std::map<uint32_t, int> uses;          // consumers left, per resolved value
    std::map<uint32_t, uint32_t> pslot, dslot;
    bool pfree[GRAPH_PSLOTS], dfree[GRAPH_DSLOTS];
    uint32_t cur = VAL_NONE, dc = VAL_NONE;
    int err = SYF_OK;
    std::vector<uint32_t> node_of;  // per instruction: the node it comes from (VAL_NONE: none)
    uint32_t node = VAL_NONE;       // the node being emitted

    explicit Compiler(const Graph& g) : G(g) {
        for (bool& f : pfree) f = true;
        for (bool& f : dfree) f = true;
    }

    uint32_t src(uint32_t id, uint32_t port) const {
        uint32_t s = VAL_NONE;
        G.input_of(id, port, &s);
        return s;
    }
    // The node whose point value flows out of `id` (a disabled point node passes its input).
    uint32_t rpoint(uint32_t id) const {
        for (int guard = 0; guard < 4096; ++guard) {
            const Node* n = G.node(id);
            if (!n || n->kind == SYF_NODE_INPUT) return VAL_START;
            if (n->enabled) return id;
            id = src(id, 0);
        }
        return VAL_START;
    }
    // The node whose distance value flows out of `id`.
    uint32_t rdist(uint32_t id) const {
        for (int guard = 0; guard < 4096; ++guard) {
            const Node* n = G.node(id);
            if (!n) return VAL_NONE;
            if (n->kind == SYF_NODE_OUTPUT) {
                id = src(id, 0);
                continue;
            }
            if (fence_kind(n->kind) || n->enabled) return id;
            id = src(id, n->kind == SYF_NODE_BOUNDED ? 1u : 0u);  // a disabled node passes its input
        }
        return VAL_NONE;
    }
    // Consumers of every value the Output needs.
    void count(uint32_t v, bool point, std::set<uint32_t>& visited) {
        if (v == VAL_START || v == VAL_NONE) return;
        uses[v]++;
        if (!visited.insert(v).second) return;
        const Node* n = G.node(v);
        if (point) {
            count(rpoint(src(v, 0)), true, visited);
        } else if (fence_kind(n->kind)) {
            count(rpoint(src(v, 0)), true, visited);
        } else if (n->kind == SYF_NODE_MATERIAL_TAG) {
            count(rdist(src(v, 0)), false, visited);
        } else if (n->kind == SYF_NODE_BOUNDED) {
            count(rpoint(src(v, 0)), true, visited);
            count(rdist(src(v, 1)), false, visited);
        } else if (combine_kind(n->kind)) {
            count(rdist(src(v, 0)), false, visited);
            count(rdist(src(v, 1)), false, visited);
        }
    }

    GOp& emit(uint32_t op, uint32_t a = 0, uint32_t b = 0, uint32_t c = 0) {
        GOp o;
        std::memset(&o, 0, sizeof o);
        o.op = op;
        o.a = a;
        o.b = b;
        o.c = c;
        prog.push_back(o);
        node_of.push_back(node);
        return prog.back();
    }
    int slot_full(uint32_t id, const char* what, uint32_t n) {
        if (err == SYF_OK)
            err = node_err(id, "the graph keeps more than %u %s values alive at once, the most this library's "
                               "interpreter holds", n, what);
        return err;
    }
    void release(uint32_t v, bool point) {
        auto& m = point ? pslot : dslot;
        auto it = m.find(v);
        if (it != m.end() && uses[v] <= 0) {
            (point ? pfree : dfree)[it->second] = true;
            m.erase(it);
        }
    }
    // Saves the current point if a later consumer still needs it.
    void spill_point() {
        if (cur == VAL_NONE || cur == VAL_START || uses[cur] <= 0 || pslot.count(cur)) return;
        for (uint32_t s = 0; s < GRAPH_PSLOTS; ++s)
            if (pfree[s]) {
                pfree[s] = false;
                pslot[cur] = s;
                emit(GOP_P_SAVE, s);
                return;
            }
        slot_full(cur, "point", GRAPH_PSLOTS);
    }
    void spill_dist() {
        if (dc == VAL_NONE || uses[dc] <= 0 || dslot.count(dc)) return;
        for (uint32_t s = 0; s < GRAPH_DSLOTS; ++s)
            if (dfree[s]) {
                dfree[s] = false;
                dslot[dc] = s;
                emit(GOP_D_SAVE, s);
                return;
            }
        slot_full(dc, "distance", GRAPH_DSLOTS);
    }

    // One enabled point node applied to the current point; `level` is the Iterate nesting
    // depth (0 outside every Iterate), `seed` the AddSeed source (0: own input, L: the entry
    // point of the enclosing Iterate at level L - 1).
    void point_op(const Node& n, uint32_t level, uint32_t seed) {
        const float* v = n.v;
        node = n.id;
        switch (n.kind) {
        case SYF_NODE_TRANSLATE: {
            GOp& o = emit(GOP_TRANSLATE);
            for (int k = 0; k < 3; ++k) o.v[k] = v[k];
            break;
        }
        case SYF_NODE_ROTATE: {
            // Quat::from_euler(XYZ, ...) as Rx Ry Rz with glam's quaternion product, in f32
            const float r = 3.14159265358979323846f / 180.0f;
            const float ha = (v[0] * r) * 0.5f, hb = (v[1] * r) * 0.5f, hc = (v[2] * r) * 0.5f;
            const float qx[4] = {std::sin(ha), 0.0f, 0.0f, std::cos(ha)};
            const float qy[4] = {0.0f, std::sin(hb), 0.0f, std::cos(hb)};
            const float qz[4] = {0.0f, 0.0f, std::sin(hc), std::cos(hc)};
            auto mul = [](const float* a, const float* b, float* o) {
                o[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
                o[1] = a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0];
                o[2] = a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3];
                o[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
            };
            float qxy[4], q[4];
            mul(qx, qy, qxy);
            mul(qxy, qz, q);
            GOp& o = emit(GOP_ROTATE);
            for (int k = 0; k < 4; ++k) o.v[k] = q[k];
            break;
        }
        case SYF_NODE_SCALE: {
            GOp& o = emit(GOP_SCALE);
            o.v[0] = v[0];
            o.v[1] = std::fabs(v[0]);
            break;
        }
        case SYF_NODE_SCALE_AXES: {
            GOp& o = emit(GOP_SCALE_AXES);
            for (int k = 0; k < 3; ++k) o.v[k] = v[k];
            o.v[3] = std::fmax(v[0], std::fmax(v[1], v[2]));
            break;
        }
        case SYF_NODE_MIRROR: {
            GOp& o = emit(GOP_MIRROR);
            normal_of(v, o.v);
            o.v[3] = v[3];
            break;
        }
        case SYF_NODE_BOX_FOLD: emit(GOP_BOX_FOLD).v[0] = v[0]; break;
        case SYF_NODE_BOX_FOLD_AXES: {
            GOp& o = emit(GOP_BOX_FOLD_AXES, n.flags & 7u);
            for (int k = 0; k < 3; ++k) o.v[k] = v[k];
            break;
        }
        case SYF_NODE_SPHERE_INVERSION: {
            GOp& o = emit(GOP_SPHERE_INVERSION);
            o.v[0] = v[0];
            o.v[1] = v[1];
            o.v[2] = v[0] * v[0];
            o.v[3] = v[1] * v[1];
            break;
        }
        case SYF_NODE_SPHERE_FOLD: {
            GOp& o = emit(GOP_SPHERE_FOLD);
            o.v[0] = v[0];
            o.v[1] = v[1];
            o.v[2] = v[0] * v[0];
            o.v[3] = v[1] * v[1];
            o.v[4] = o.v[3] / o.v[2];
            break;
        }
        case SYF_NODE_ABS: emit(GOP_ABS, n.flags & 7u); break;
        case SYF_NODE_MENGER_FOLD: {
            GOp& o = emit(GOP_MENGER_FOLD);
            const float s1 = v[0] - 1.0f;
            o.v[0] = v[0];
            for (int k = 0; k < 3; ++k) {
                o.v[1 + k] = v[1 + k];
                o.v[4 + k] = v[1 + k] * s1;
            }
            o.v[7] = -0.5f * v[3] * s1;
            break;
        }
        case SYF_NODE_KALEIDO: {
            GOp& o = emit(GOP_KALEIDO);
            o.v[0] = 6.28318530717958647692f / float(std::max(n.u[0], 1u));
            o.v[1] = v[0] * (3.14159265358979323846f / 180.0f);
            break;
        }
        case SYF_NODE_POWER: {
            const float p = v[0];
            const bool integral = p >= 1.0f && p <= 16.0f && p == std::floor(p);
            GOp& o = emit(GOP_POWER, integral ? uint32_t(p) : 0u);
            o.v[0] = p;
            break;
        }
        case SYF_NODE_ADD_ORIGIN: emit(GOP_ADD_ORIGIN); break;
        case SYF_NODE_ADD_SEED: {
            GOp& o = emit(GOP_ADD_SEED, seed);
            for (int k = 0; k < 3; ++k) o.v[k] = v[k];
            o.v[3] = std::fmax(std::fabs(v[0]), std::fmax(std::fabs(v[1]), std::fabs(v[2])));
            break;
        }
        case SYF_NODE_ADD_CONSTANT: {
            GOp& o = emit(GOP_ADD_CONSTANT);
            for (int k = 0; k < 6; ++k) o.v[k] = v[k];
            o.v[6] = std::fmax(std::fabs(v[0]), std::fmax(std::fabs(v[1]), std::fabs(v[2])));
            break;
        }
        case SYF_NODE_TWIST: emit(GOP_TWIST).v[0] = v[0]; break;
        case SYF_NODE_BEND: emit(GOP_BEND).v[0] = v[0]; break;
        case SYF_NODE_TILE: {
            GOp& o = emit(GOP_TILE);
            for (int k = 0; k < 3; ++k) o.v[k] = v[k];
            break;
        }
        case SYF_NODE_REPEAT_FINITE: {
            GOp& o = emit(GOP_REPEAT_FINITE, n.u[0], n.u[1], n.u[2]);
            for (int k = 0; k < 3; ++k) o.v[k] = v[k];
            break;
        }
        case SYF_NODE_TRAP: emit(GOP_TRAP, std::min(n.u[0], 3u), n.u[1], n.u[2]); break;
        case SYF_NODE_ITERATE: {
            const size_t begin = prog.size();
            emit(GOP_ITER_BEGIN, level, n.u[0]).v[0] = v[0];
            const uint32_t body_start = uint32_t(prog.size());
            for (uint32_t b : n.body) {
                const Node* m = G.node(b);
                if (m && m->enabled) point_op(*m, level + 1, level + 1);
            }
            const uint32_t end = uint32_t(prog.size());
            GOp& o = emit(GOP_ITER_END, level, n.u[0], body_start);
            o.v[0] = v[0];
            prog[begin].c = end;
            break;
        }
        default: break;
        }
    }
    // try_normalize().unwrap_or(Y) in f32
    static void normal_of(const float* v, float* out) {
        const float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        const float rcp = 1.0f / l;
        if (std::isfinite(rcp) && rcp > 0.0f) {
            for (int k = 0; k < 3; ++k) out[k] = v[k] * rcp;
        } else {
            out[0] = 0.0f;
            out[1] = 1.0f;
            out[2] = 0.0f;
        }
    }

    // Makes the current point hold value v.
    void gen_point(uint32_t v) {
        if (err != SYF_OK || cur == v) return;
        if (v == VAL_START) {
            spill_point();
            emit(GOP_P_START);
            cur = VAL_START;
            return;
        }
        auto it = pslot.find(v);
        if (it != pslot.end()) {
            spill_point();
            emit(GOP_P_LOAD, it->second);
            cur = v;
            return;
        }
        const Node& n = *G.node(v);
        const uint32_t u = rpoint(src(v, 0));
        gen_point(u);
        if (u != VAL_START) uses[u]--;
        spill_point();  // the operation below replaces the current point
        if (u != VAL_START) release(u, true);
        point_op(n, 0, 0);
        cur = v;
    }
    void use_point(uint32_t u) {
        if (u == VAL_START) return;
        uses[u]--;
        release(u, true);
    }
    void use_dist(uint32_t d) {
        uses[d]--;
        release(d, false);
    }

    // Makes the current distance hold value v.
    void gen_dist(uint32_t v) {
        if (err != SYF_OK || dc == v) return;
        if (v == VAL_NONE) {
            err = graph_err("a distance input is not connected");
            return;
        }
        auto it = dslot.find(v);
        if (it != dslot.end()) {
            spill_dist();
            emit(GOP_D_LOAD, it->second);
            dc = v;
            return;
        }
        const Node& n = *G.node(v);
        if (fence_kind(n.kind)) {
            const uint32_t u = rpoint(src(v, 0));
            gen_point(u);
            use_point(u);
            spill_dist();
            fence_op(n);
        } else if (n.kind == SYF_NODE_MATERIAL_TAG) {
            const uint32_t s = rdist(src(v, 0));
            gen_dist(s);
            uses[s]--;
            spill_dist();
            release(s, false);
            emit(GOP_MATTAG, n.u[0]);
        } else if (n.kind == SYF_NODE_BOUNDED) {
            const uint32_t s = rdist(src(v, 1)), u = rpoint(src(v, 0));
            gen_dist(s);
            gen_point(u);
            uses[s]--;
            spill_dist();
            release(s, false);
            use_point(u);
            GOp& o = emit(GOP_BOUNDED);
            for (int k = 0; k < 3; ++k) o.v[k] = n.v[k];
        } else if (combine_kind(n.kind)) {
            const uint32_t a = rdist(src(v, 0)), b = rdist(src(v, 1));
            gen_dist(a);
            gen_dist(b);  // saves a first when it replaces the current distance
            if (err != SYF_OK) return;
            const uint32_t sa = (dc == a) ? GRAPH_ACC : dslot.at(a);
            const uint32_t sb = (dc == b) ? GRAPH_ACC : dslot.at(b);
            uses[a]--;
            uses[b]--;
            spill_dist();  // b, if a later consumer needs it
            GOp& o = emit(GOP_COMBINE, n.kind, sa, sb);
            o.v[0] = n.v[0];
            release(a, false);
            release(b, false);
        } else if (n.kind == SYF_NODE_INSTANCES) {
            err = node_err(n.id, "“Independent Copies” (Instances) is not evaluated by this library");
            return;
        }
        dc = v;
    }
    void fence_op(const Node& n) {
        node = n.id;
        GOp& o = emit(GOP_FENCE, n.enabled ? n.kind : SYF_NODE_KINDS, n.kind == SYF_NODE_ROCK ? n.u[0] : 0u);
        for (int k = 0; k < 6; ++k) o.v[k] = n.v[k];
        if (n.kind == SYF_NODE_PLANE) normal_of(n.v, o.v);
    }
};

// Checks the invariants graph_run relies on: a known opcode everywhere, slot indices below the
// slot counts, ITER_BEGIN / ITER_END pairs properly nested within GRAPH_LEVELS and pointing at
// each other, AddSeed levels inside their loops, and GOP_END last.  A failure is an error of
// this library, reported as such.
int verify(const std::vector<GOp>& prog) {
    std::vector<uint32_t> open;  // indices of the enclosing ITER_BEGINs
    for (uint32_t i = 0; i < prog.size(); ++i) {
        const GOp& o = prog[i];
        bool ok = o.op < GOP_COUNT;
        switch (o.op) {
        case GOP_P_SAVE:
        case GOP_P_LOAD: ok = ok && o.a < GRAPH_PSLOTS && open.empty(); break;
        case GOP_D_SAVE:
        case GOP_D_LOAD: ok = ok && o.a < GRAPH_DSLOTS && open.empty(); break;
        case GOP_COMBINE:
            ok = ok && (o.b == GRAPH_ACC || o.b < GRAPH_DSLOTS) && (o.c == GRAPH_ACC || o.c < GRAPH_DSLOTS) && open.empty();
            break;
        case GOP_ITER_BEGIN:
            ok = ok && o.a == open.size() && o.a < GRAPH_LEVELS && o.c > i && o.c < prog.size() &&
                 prog[o.c].op == GOP_ITER_END && prog[o.c].a == o.a && prog[o.c].c == i + 1;
            open.push_back(i);
            break;
        case GOP_ITER_END:
            ok = ok && !open.empty() && prog[open.back()].c == i && o.a + 1 == open.size();
            if (!open.empty()) open.pop_back();
            break;
        case GOP_ADD_SEED: ok = ok && o.a <= open.size(); break;
        case GOP_TRAP: ok = ok && o.a < 4; break;
        case GOP_END: ok = ok && i + 1 == prog.size() && open.empty(); break;
        default: break;
        }
        if (!ok) return fail(SYF_E_INVALID, "graph: internal error: compiled instruction %u (opcode %u) is malformed", i, o.op);
    }
    if (prog.empty() || prog.back().op != GOP_END) return fail(SYF_E_INVALID, "graph: internal error: program without end");
    return SYF_OK;
}

// The classic family a compiled graph is, if any: the program of the bundled Mandelbulb,
// Mandelbox, Menger or sphere graph (graphs/*.ron) whatever its parameters, every node enabled,
// the trap in slot 0 measuring the norm at its minimum, the Mandelbox's norm surface without
// inset.  Those graphs run on the classic families' kernels, which compute the same values (the
// same formulas in the same order; bit-identical in exact arithmetic) and also have the fast FP32
// march.  Fills out.classic from the nodes' own parameters.
void classic_of(const Graph& G, Compiled& out) {
    const std::vector<GOp>& p = out.prog;
    auto node = [&](size_t i) -> const Node& { return *G.node(out.node_of[i]); };
    auto ops = [&](std::initializer_list<uint32_t> want) {
        if (p.size() != want.size()) return false;
        size_t i = 0;
        for (uint32_t w : want)
            if (p[i++].op != w) return false;
        return true;
    };
    auto trap_ok = [&](size_t i) { return p[i].a == 0u && p[i].b == SYF_MEASURE_NORM && p[i].c == SYF_TRAP_MINIMUM; };
    syf_fractal f;
    std::memset(&f, 0, sizeof f);
    f.safety = G.safety;
    f.bounding_radius_hint = G.hint;
    f.trap_scale = out.trap_scales[0];
    f.plane_normal[1] = 1.0f;
    uint32_t fam = SYF_FAMILY_GRAPH;
    if (ops({GOP_P_START, GOP_ITER_BEGIN, GOP_TRAP, GOP_POWER, GOP_ADD_ORIGIN, GOP_ITER_END, GOP_FENCE, GOP_END}) &&
        trap_ok(2) && p[6].a == SYF_NODE_BULB_FENCE) {
        fam = SYF_FAMILY_MANDELBULB;
        f.power = node(3).v[0];
    } else if (ops({GOP_P_START, GOP_ITER_BEGIN, GOP_BOX_FOLD, GOP_TRAP, GOP_SPHERE_FOLD, GOP_SCALE, GOP_ADD_ORIGIN,
                    GOP_ITER_END, GOP_FENCE, GOP_END}) &&
               trap_ok(3) && p[8].a == SYF_NODE_NORM_FENCE && node(8).v[0] == 0.0f) {
        fam = SYF_FAMILY_MANDELBOX;
        f.fold_limit = node(2).v[0];
        f.min_radius = node(4).v[0];
        f.fixed_radius = node(4).v[1];
        f.scale = node(5).v[0];
    } else if (ops({GOP_P_START, GOP_ITER_BEGIN, GOP_MENGER_FOLD, GOP_TRAP, GOP_ITER_END, GOP_FENCE, GOP_END}) &&
               trap_ok(3) && p[5].a == SYF_NODE_BOX) {
        fam = SYF_FAMILY_MENGER;
        f.scale = node(2).v[0];
        for (int k = 0; k < 3; ++k) {
            f.offset[k] = node(2).v[1 + k];
            f.half[k] = node(5).v[k];
        }
    } else if (ops({GOP_P_START, GOP_TRAP, GOP_FENCE, GOP_D_SAVE, GOP_FENCE, GOP_COMBINE, GOP_END}) && trap_ok(1) &&
               p[2].a == SYF_NODE_SPHERE && p[4].a == SYF_NODE_PLANE && p[5].a == SYF_NODE_UNION &&
               p[3].a == p[5].b && p[5].c == GRAPH_ACC) {
        fam = SYF_FAMILY_SPHERE;
        f.radius = node(2).v[0];
        for (int k = 0; k < 3; ++k) f.plane_normal[k] = node(4).v[k];
        f.plane_offset = node(4).v[3];
    }
    if (fam == SYF_FAMILY_GRAPH) return;
    if (fam != SYF_FAMILY_SPHERE) {
        f.iterations = p[1].b;
        f.bailout = p[1].v[0];
    }
    f.family = fam;
    out.classic = f;
    out.classic_family = fam;
}

// Validates, derives and compiles a graph.
int compile(const syf_graph& cg, Compiled& out) {
    Graph G;
    if (int rc = load(cg, G); rc != SYF_OK) return rc;
    if (int rc = validate(G); rc != SYF_OK) return rc;
    for (const Node& n : G.nodes)
        if (n.kind == SYF_NODE_INSTANCES && n.enabled) {
            // only a node the Output reaches matters
            const std::vector<uint32_t> act = G.active();
            if (std::find(act.begin(), act.end(), n.id) != act.end())
                return node_err(n.id, "“Independent Copies” (Instances) is not evaluated by this library");
        }
    std::vector<uint32_t> order;
    G.topo_order(&order);
    for (float& s : out.trap_scales) s = 1.0f;
    for (uint32_t id : order) G.trap_walk(id, 0, out.trap_scales);
    out.iteration_product = G.iteration_product();
    out.step_scale = G.step_scale();
    out.hint = G.hint;
    Compiler C(G);
    const uint32_t root = C.rdist(G.output()->id);
    std::set<uint32_t> visited;
    C.count(root, false, visited);
    C.uses[root]++;  // the program's result
    C.gen_dist(root);
    if (C.err != SYF_OK) return C.err;
    C.emit(GOP_END);
    if (C.prog.size() > GRAPH_MAX_OPS)
        return graph_err("the compiled graph has %zu instructions, above this library's %u", C.prog.size(), GRAPH_MAX_OPS);
    if (int rc = verify(C.prog); rc != SYF_OK) return rc;
    out.prog = std::move(C.prog);
    out.node_of = std::move(C.node_of);
    classic_of(G, out);
    return SYF_OK;
}

// ---------------------------------------------------------------- specialised topologies

// The evaluator of one compiled program as straight-line device code: the program's structure
// (instruction sequence, loop nesting, slots, fence and combination kinds, trap slots,
// measures and modes, integer powers, axis masks, AddSeed levels) is written into the code;
// every value the structure does not fix (the node parameters, loop counts and bailouts,
// material indices) is read from the program in device memory at run time, so graphs of one
// topology share their kernels whatever their parameters, as the reference's pipeline cache
// shares a shader per topology key (NodeKind::topology_key, mod.rs:371-398).  The operations
// are graph_device.hpp's, the interpreter's own: a specialised graph computes what the
// interpreter computes.
const char* const GOP_NAME[GOP_COUNT] = {
    "GOP_END", "GOP_P_START", "GOP_P_SAVE", "GOP_P_LOAD", "GOP_ITER_BEGIN", "GOP_ITER_END", "GOP_TRANSLATE",
    "GOP_ROTATE", "GOP_SCALE", "GOP_SCALE_AXES", "GOP_MIRROR", "GOP_BOX_FOLD", "GOP_BOX_FOLD_AXES",
    "GOP_SPHERE_INVERSION", "GOP_SPHERE_FOLD", "GOP_ABS", "GOP_MENGER_FOLD", "GOP_KALEIDO", "GOP_POWER",
    "GOP_ADD_ORIGIN", "GOP_ADD_SEED", "GOP_ADD_CONSTANT", "GOP_TWIST", "GOP_BEND", "GOP_TILE", "GOP_REPEAT_FINITE",
    "GOP_TRAP", "GOP_FENCE", "GOP_D_SAVE", "GOP_D_LOAD", "GOP_MATTAG", "GOP_BOUNDED", "GOP_COMBINE"};
static_assert(GOP_COMBINE + 1 == GOP_COUNT, "a name per opcode");

std::string jit_source(const std::vector<GOp>& prog) {
    std::string o;
    char buf[512];
    auto line = [&](const char* fmt, auto... a) {
        std::snprintf(buf, sizeof buf, fmt, a...);
        o += buf;
        o += '\n';
    };
    uint32_t np = 0, nd = 0;
    for (const GOp& g : prog) {
        if (g.op == GOP_P_SAVE || g.op == GOP_P_LOAD) np = std::max(np, g.a + 1);
        if (g.op == GOP_D_SAVE || g.op == GOP_D_LOAD) nd = std::max(nd, g.a + 1);
        if (g.op == GOP_COMBINE) {
            if (g.b != GRAPH_ACC) nd = std::max(nd, g.b + 1);
            if (g.c != GRAPH_ACC) nd = std::max(nd, g.c + 1);
        }
    }
    line("#include \"trace_kernels.hpp\"");
    line("#include \"pathtrace_kernels.hpp\"");
    line("namespace syf {");
    line("template <class T> __device__ __forceinline__ GRaw<T> graph_jit(const GOp* prog_in, V3<T> c) {");
    line("const GOp* __restrict__ prog = g_uniform(prog_in);");
    line("const T INF = g_inf<T>();");
    line("GPoint<T> cur{c, T(1), 0u, INF, INF, INF, INF};");
    line("GDist<T> dc{INF, 0u, 0u, INF, INF, INF, INF};");
    line("uint32_t runs = 0u;");
    for (uint32_t k = 0; k < np; ++k) line("GPoint<T> ps%u;", k);
    for (uint32_t k = 0; k < nd; ++k) line("GDist<T> ds%u;", k);
    for (uint32_t i = 0; i < prog.size(); ++i) {
        const GOp& g = prog[i];
        switch (g.op) {
        case GOP_END: line("return GRaw<T>{dc, runs};"); break;
        case GOP_P_START: line("cur = GPoint<T>{c, T(1), 0u, INF, INF, INF, INF};"); break;
        case GOP_P_SAVE: line("ps%u = cur;", g.a); break;
        case GOP_P_LOAD: line("cur = ps%u;", g.a); break;
        case GOP_D_SAVE: line("ds%u = dc;", g.a); break;
        case GOP_D_LOAD: line("dc = ds%u;", g.a); break;
        case GOP_ITER_BEGIN:
            line("{ const V3<T> sp%u = cur.p; const T sd%u = cur.dr;", g.a, g.a);
            line("const uint32_t n%u = prog[%u].b; const T bail%u = T(prog[%u].v[0]);", i, i, i, i);
            line("for (uint32_t k%u = 0u; k%u < n%u; ++k%u) { if (len3(cur.p) > bail%u) break;", i, i, i, i, i);
            break;
        case GOP_ITER_END: line("cur.iter += 1u; runs += 1u; } }"); break;
        case GOP_ADD_SEED:
            if (g.a == 0u) line("g_point<T>(%s, 0u, 0u, 0u, prog[%u], cur, c, cur.p, cur.dr);", GOP_NAME[g.op], i);
            else line("g_point<T>(%s, 0u, 0u, 0u, prog[%u], cur, c, sp%u, sd%u);", GOP_NAME[g.op], i, g.a - 1, g.a - 1);
            break;
        case GOP_FENCE: line("dc = g_fence_state<T>(%uu, prog[%u], cur);", g.a, i); break;
        case GOP_MATTAG: line("dc.mat = prog[%u].a;", i); break;
        case GOP_BOUNDED: line("g_bounded<T>(prog[%u], cur, dc);", i); break;
        case GOP_COMBINE: {
            char a[16], b[16];
            std::snprintf(a, sizeof a, g.b == GRAPH_ACC ? "dc" : "ds%u", g.b);
            std::snprintf(b, sizeof b, g.c == GRAPH_ACC ? "dc" : "ds%u", g.c);
            line("dc = g_combine<T>(%uu, T(prog[%u].v[0]), GDist<T>(%s), GDist<T>(%s));", g.a, i, a, b);
            break;
        }
        case GOP_REPEAT_FINITE:  // the counts are parameters: read from the program
            line("g_point<T>(%s, 0u, 0u, 0u, prog[%u], cur, c, cur.p, cur.dr);", GOP_NAME[g.op], i);
            break;
        default:  // point operations: a, b, c (trap slot / measure / mode, power, masks) are structure
            line("g_point<T>(%s, %uu, %uu, %uu, prog[%u], cur, c, cur.p, cur.dr);", GOP_NAME[g.op], g.a, g.b, g.c, i);
            break;
        }
    }
    line("return GRaw<T>{dc, runs};");
    line("}");
    line("}  // namespace syf");
    return o;
}

const char* const JIT_NAMES[JIT_KERNELS] = {
    "syf::k_primary<7u>",   "syf::k_primary_exact<7u>", "syf::k_hits_tail<7u>", "syf::k_de_probe<7u>",
    "syf::k_render_tile<7u>", "syf::k_wf_intersect<7u>", "syf::k_wf_shade<7u>",  "syf::k_wf_commit<7u>",
    "syf::k_render_tile_full<7u>"};
static_assert(FAM_JIT == 7u, "the run-time kernel names carry the family");

struct JitModule {
    hipModule_t module[JIT_KERNELS] = {};
    hipFunction_t fn[JIT_KERNELS] = {};
    double compile_ms = 0.0;
};

std::mutex g_jit_mutex;
std::map<std::pair<int, std::string>, JitModule*> g_jit_cache;  // (device, source): loaded for the process

// The directory of the library's device sources (csrc/ beside libsyfrack_hip.so), or
// SYF_CSRC_DIR.
std::string csrc_dir() {
    if (const char* e = std::getenv("SYF_CSRC_DIR")) return e;
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&syf_graph_info), &info) && info.dli_fname)
        return (std::filesystem::path(info.dli_fname).parent_path() / "csrc").string();
    return "csrc";
}
std::string rocm_dir() {
    if (const char* e = std::getenv("ROCM_PATH")) return e;
    return "/opt/rocm";
}
// The compiler's own headers (stddef.h, stdint.h), which run-time compilation does not add by itself.
std::string clang_include_dir(const std::string& rocm) {
    std::error_code ec;
    const std::filesystem::path base = std::filesystem::path(rocm) / "lib/llvm/lib/clang";
    std::string best;
    for (const auto& d : std::filesystem::directory_iterator(base, ec))
        if (std::filesystem::exists(d.path() / "include/stddef.h")) best = (d.path() / "include").string();
    return best;
}

// One kernel of a topology compiled by hiprtc: its code object and lowered name, or the error.
struct JitPiece {
    std::vector<char> code;
    std::string lowered, error;
};
void jit_compile_one(const std::string& src, const char* name, const std::vector<std::string>& options, JitPiece* out) {
    std::vector<const char*> opts;
    for (const std::string& o : options) opts.push_back(o.c_str());
    hiprtcProgram rp = nullptr;
    if (hiprtcCreateProgram(&rp, src.c_str(), "graph_topology.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        out->error = "hiprtcCreateProgram failed";
        return;
    }
    hiprtcAddNameExpression(rp, name);
    const hiprtcResult cr = hiprtcCompileProgram(rp, int(opts.size()), opts.data());
    if (cr != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(rp, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(rp, &log[0]);
        out->error = std::string(hiprtcGetErrorString(cr)) + ": " + log.substr(0, 600);
        hiprtcDestroyProgram(&rp);
        return;
    }
    size_t code_n = 0;
    hiprtcGetCodeSize(rp, &code_n);
    out->code.resize(code_n);
    hiprtcGetCode(rp, out->code.data());
    const char* lowered = nullptr;
    if (hiprtcGetLoweredName(rp, name, &lowered) == HIPRTC_SUCCESS && lowered) out->lowered = lowered;
    else out->error = std::string("no compiled kernel ") + name;
    hiprtcDestroyProgram(&rp);
}

// The kernels of a program's topology on the current device: from the process cache, or
// compiled (hiprtc, one program per kernel, all at once on host threads) and loaded.
int jit_get(int device, const std::vector<GOp>& prog, const JitModule** out) {
    const std::string src = jit_source(prog);
    std::lock_guard<std::mutex> lock(g_jit_mutex);
    auto it = g_jit_cache.find({device, src});
    if (it != g_jit_cache.end()) {
        *out = it->second;
        return SYF_OK;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const std::string rocm = rocm_dir(), clang_inc = clang_include_dir(rocm);
    if (clang_inc.empty())
        return fail(SYF_E_DEVICE, "graph: the compiler headers (stddef.h) were not found under %s/lib/llvm/lib/clang",
                    rocm.c_str());
    const std::vector<std::string> options = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                                              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-w",
                                              "-I" + csrc_dir(), "-I" + rocm + "/include", "-I" + clang_inc};
    JitPiece pieces[JIT_KERNELS];
    std::vector<std::thread> threads;
    for (uint32_t k = 0; k < JIT_KERNELS; ++k)
        threads.emplace_back(jit_compile_one, std::cref(src), JIT_NAMES[k], std::cref(options), &pieces[k]);
    for (std::thread& t : threads) t.join();
    for (uint32_t k = 0; k < JIT_KERNELS; ++k)
        if (!pieces[k].error.empty())
            return fail(SYF_E_DEVICE, "graph: run-time compilation of %s for the topology failed: %s", JIT_NAMES[k],
                        pieces[k].error.c_str());
    JitModule* m = new JitModule();
    hipError_t e = hipSuccess;
    for (uint32_t k = 0; k < JIT_KERNELS && e == hipSuccess; ++k) {
        e = hipModuleLoadData(&m->module[k], pieces[k].code.data());
        if (e == hipSuccess) e = hipModuleGetFunction(&m->fn[k], m->module[k], pieces[k].lowered.c_str());
    }
    if (e != hipSuccess) {
        for (hipModule_t mod : m->module)
            if (mod) (void)hipModuleUnload(mod);
        delete m;
        return hip_fail(e, "graph: loading the topology's kernels");
    }
    m->compile_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    g_jit_cache[{device, src}] = m;
    *out = m;
    return SYF_OK;
}

}  // namespace

int graph_lower_interpreted(const syf_job_desc& desc, const syf_graph& graph, JobP& P, std::vector<GOp>& prog,
                            std::vector<uint32_t>& node_of) {
    Compiled c;
    if (int rc = compile(graph, c); rc != SYF_OK) return rc;
    FractalP F;
    std::memset(&F, 0, sizeof F);
    F.family = SYF_FAMILY_GRAPH;
    F.iteration_product = c.iteration_product;
    F.safety = c.step_scale;
    for (int i = 0; i < 4; ++i) F.gtrap_scale[i] = c.trap_scales[i];
    F.trap_scale = c.trap_scales[0];
    syf_job_desc d = desc;
    d.fractal.family = SYF_FAMILY_GRAPH;
    d.fractal.bounding_radius_hint = c.hint;
    if (int rc = lower_job(d, P, &F); rc != SYF_OK) return rc;
    prog = std::move(c.prog);
    node_of = std::move(c.node_of);
    return SYF_OK;
}

hipError_t jit_launch_args(const void* jit, uint32_t which, dim3 grid, dim3 block, size_t shmem, hipStream_t s,
                           void** args) {
    const JitModule* m = static_cast<const JitModule*>(jit);
    if (!m || which >= JIT_KERNELS || !m->fn[which]) return hipErrorInvalidDeviceFunction;
    return hipModuleLaunchKernel(m->fn[which], grid.x, grid.y, grid.z, block.x, block.y, block.z, uint32_t(shmem), s,
                                 args, nullptr);
}

}  // namespace syf

using namespace syf;

extern "C" {

int syf_graph_info(const syf_graph* graph, float* step_scale, uint32_t* iteration_product, float* trap_scales4,
                   uint32_t* kernel_family) {
    clear_error();
    if (!graph) return fail(SYF_E_INVALID, "graph must not be null");
    Compiled c;
    if (int rc = compile(*graph, c); rc != SYF_OK) return rc;
    if (kernel_family) *kernel_family = c.classic_family;
    if (step_scale) *step_scale = c.step_scale;
    if (iteration_product) *iteration_product = c.iteration_product;
    if (trap_scales4)
        for (int i = 0; i < 4; ++i) trap_scales4[i] = c.trap_scales[i];
    return SYF_OK;
}

int syf_graph_source(const syf_graph* graph, char* buf, size_t cap, size_t* len) {
    clear_error();
    if (!graph || !len) return fail(SYF_E_INVALID, "graph and len must not be null");
    Compiled c;
    if (int rc = compile(*graph, c); rc != SYF_OK) return rc;
    const std::string src = jit_source(c.prog);
    *len = src.size() + 1;
    if (buf && cap > 0) {
        const size_t n = std::min(cap - 1, src.size());
        std::memcpy(buf, src.data(), n);
        buf[n] = '\0';
    }
    return SYF_OK;
}

int syf_job_create_graph(syf_ctx* ctx, const syf_job_desc* desc, const syf_graph* graph, uint32_t mode,
                         syf_job** out) {
    clear_error();
    if (!ctx || !desc || !graph || !out) return fail(SYF_E_INVALID, "ctx, desc, graph and out must not be null");
    if (mode != SYF_GRAPH_SPECIALISED && mode != SYF_GRAPH_INTERPRETED)
        return fail(SYF_E_INVALID, "mode must be SYF_GRAPH_SPECIALISED (0) or SYF_GRAPH_INTERPRETED (1), got %u", mode);
    if (desc->fractal.family != SYF_FAMILY_GRAPH)
        return fail(SYF_E_INVALID, "fractal.family must be SYF_FAMILY_GRAPH (%u) for a graph job, got %u", SYF_FAMILY_GRAPH,
                    desc->fractal.family);
    Compiled c;
    if (int rc = compile(*graph, c); rc != SYF_OK) return rc;
    if (mode == SYF_GRAPH_SPECIALISED && c.classic_family != SYF_FAMILY_GRAPH) {
        // a classic graph: the classic family's kernels
        syf_job_desc d = *desc;
        d.fractal = c.classic;
        return syf_job_create(ctx, &d, out);
    }
    FractalP F;
    std::memset(&F, 0, sizeof F);
    F.family = SYF_FAMILY_GRAPH;
    F.iteration_product = c.iteration_product;
    F.safety = c.step_scale;  // GpuFractal p0.w: safety / Lipschitz
    for (int i = 0; i < 4; ++i) F.gtrap_scale[i] = c.trap_scales[i];
    F.trap_scale = c.trap_scales[0];
    syf_job_desc d = *desc;
    d.fractal.bounding_radius_hint = c.hint;  // the ray extent follows the graph's hint (uniforms.rs:781-782)
    JobP P;
    if (int rc = lower_job(d, P, &F); rc != SYF_OK) return rc;
    if (hipError_t e = hipSetDevice(ctx->device); e != hipSuccess) return hip_fail(e, "hipSetDevice");
    GOp* dprog = nullptr;
    const size_t bytes = c.prog.size() * sizeof(GOp);
    hipError_t e = hipMalloc(&dprog, bytes);
    if (e == hipSuccess) e = hipMemcpy(dprog, c.prog.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (dprog) (void)hipFree(dprog);
        return hip_fail(e, "graph program upload");
    }
    P.f.gprog = dprog;
    if (mode == SYF_GRAPH_SPECIALISED) {
        const JitModule* m = nullptr;
        if (int rc = jit_get(ctx->device, c.prog, &m); rc != SYF_OK) {
            (void)hipFree(dprog);
            return rc;
        }
        P.f.gjit = m;
    }
    syf_job* j = nullptr;
    if (int rc = job_alloc(ctx, d, P, &j); rc != SYF_OK) {
        (void)hipFree(dprog);
        return rc;
    }
    j->gprog = dprog;
    j->gprog_host.resize(bytes);
    std::memcpy(j->gprog_host.data(), c.prog.data(), bytes);
    *out = j;
    return SYF_OK;
}

}  // extern "C"
