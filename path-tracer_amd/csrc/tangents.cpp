// tangents.cpp - per-corner tangent frames of triangle meshes by Mikkelsen's
// MikkTSpace method ("Simulation of Wrinkled Surfaces Revisited", 2008), the
// method behind the bevy_mikktspace crate the reference calls for glTF
// primitives without usable tangents (world.rs:1250-1266; Cargo.toml:
// bevy_mikktspace 1.0.0 with corrected edge sorting).
//
// The crate is not part of the reference tree; this is a restatement of the
// published algorithm with its default 180-degree angular threshold:
//   1. weld corners whose position, normal and texcoord are equal;
//   2. triangles with two equal corner positions are degenerate and set aside;
//   3. per triangle: the texture-space derivative directions (os, ot), their
//      magnitudes per unit uv area, and the orientation (sign of the uv area);
//      triangles with zero uv area or zero derivatives "group with anything";
//   4. neighbours across shared welded edges (edges sorted by (min, max, face));
//   5. groups: the triangles around each welded vertex reachable through
//      neighbours with the same orientation (depth-first, left edge first);
//   6. each group's frame: the angle-weighted sum of the members' derivatives
//      projected into the vertex's tangent plane, normalised (members in
//      ascending triangle order);
//   7. corners of degenerate triangles copy the frame of the first good
//      corner with the same welded vertex.
// Arithmetic follows the method's float formulation (dot = x*x' + y*y' + z*z'
// left to right, normalise = v * (1 / sqrt(dot)), "non-zero" = |x| > FLT_MIN).
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "host_parallel.h"
#include "tangents.h"

namespace ptcore {

namespace {

struct V {
    float x, y, z;
};
inline V vadd(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V vscale(float s, V v) { return {s * v.x, s * v.y, s * v.z}; }
inline float vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float vlen(V v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline V vnorm(V v) { return vscale(1.0f / vlen(v), v); }
inline bool nz(float f) { return std::fabs(f) > FLT_MIN; }
inline bool vnz(V v) { return nz(v.x) || nz(v.y) || nz(v.z); }
inline bool veq(V a, V b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

constexpr int MARK_DEGENERATE = 1, GROUP_WITH_ANY = 4, ORIENT_PRESERVING = 8;

struct Tri {
    int neighbors[3] = {-1, -1, -1};
    int group[3] = {-1, -1, -1};
    V os{0, 0, 0}, ot{0, 0, 0};
    float mag_s = 0, mag_t = 0;
    int face = 0;  // original face
    int flag = 0;
};

struct Group {
    int rep;  // welded vertex
    bool orient;
    std::vector<int> faces;
};

struct Frame {
    V os{1, 0, 0};
    V ot{0, 1, 0};
    bool orient = false;
};

struct Mesh {
    const MeshView& m;
    V pos(int idx) const {  // idx = 4 * face + corner
        const float* p = &m.positions[3 * (size_t)m.indices[3 * (size_t)(idx >> 2) + (idx & 3)]];
        return {p[0], p[1], p[2]};
    }
    V nrm(int idx) const {
        const float* p = &m.normals[3 * (size_t)m.indices[3 * (size_t)(idx >> 2) + (idx & 3)]];
        return {p[0], p[1], p[2]};
    }
    V tex(int idx) const {
        const float* p = &m.texcoords[2 * (size_t)m.indices[3 * (size_t)(idx >> 2) + (idx & 3)]];
        return {p[0], p[1], 1.0f};
    }
};

struct KeyHash {
    size_t operator()(const std::array<uint32_t, 8>& k) const {
        uint64_t h = 1469598103934665603ull;
        for (uint32_t v : k) h = (h ^ v) * 1099511628211ull;
        return (size_t)h;
    }
};

uint32_t key_bits(float f) {
    if (f == 0.0f) f = 0.0f;  // -0 welds with +0, as float equality does
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

}  // namespace

bool generate_tangents(const MeshView& mv, std::vector<float>& out) {
    const int nf = (int)(mv.index_count / 3);
    out.assign((size_t)nf * 3 * 4, 0.0f);
    if (nf == 0) return true;
    const Mesh M{mv};
    // 1. weld: representative = first corner (in face order) with equal data;
    //    NaN never compares equal, so such corners stay unique
    std::vector<int> list((size_t)nf * 3);
    {
        // Corners that share a vertex index share their data, so the weld
        // works per vertex: each vertex's first corner in face order, its
        // key, and then, over the vertices with equal keys, the least first
        // corner, which is the representative. A vertex with a NaN never
        // welds. Keys are split among the threads by hash, so each key is
        // resolved by one thread, and a least-corner rule does not depend on
        // the order or the thread count.
        const size_t nc = (size_t)nf * 3;
        uint32_t nv = 0;
        for (size_t j = 0; j < nc; ++j) nv = std::max(nv, mv.indices[j] + 1u);
        constexpr int NONE = std::numeric_limits<int>::max();
        std::vector<int> first(nv, NONE);  // vertex -> first corner slot (3 * face + corner)
        for (size_t slot = nc; slot-- > 0;) first[mv.indices[slot]] = (int)slot;
        std::vector<std::array<uint32_t, 8>> keys(nv);
        std::vector<uint64_t> hashes(nv);
        std::vector<uint8_t> has_nan(nv, 0);
        const unsigned threads = nv >= 65536 ? host_threads() : 1u;
        const KeyHash hash;
        run_slices(nv, threads, [&](size_t a, size_t e) {
            for (size_t v = a; v < e; ++v) {
                if (first[v] == NONE) continue;
                const float* p = &mv.positions[3 * v];
                const float* n = &mv.normals[3 * v];
                const float* t = &mv.texcoords[2 * v];
                const float vals[8] = {p[0], p[1], p[2], n[0], n[1], n[2], t[0], t[1]};
                for (int c = 0; c < 8; ++c) {
                    has_nan[v] |= std::isnan(vals[c]) ? 1 : 0;
                    keys[v][c] = key_bits(vals[c]);
                }
                hashes[v] = hash(keys[v]);
            }
        });
        std::vector<int> rep(nv, NONE);  // vertex -> least first corner over its key
        run_parallel(threads, [&](unsigned part) {
            auto mine = [&](size_t v) {
                return first[v] != NONE && !has_nan[v] && (uint32_t)hashes[v] % threads == part;
            };
            size_t count = 0;
            for (size_t v = 0; v < nv; ++v) count += mine(v);
            size_t cap = 16;
            while (cap < 2 * count) cap <<= 1;
            std::vector<int> owner(cap, -1), least(cap);
            auto find = [&](size_t v) {
                size_t h = (size_t)(hashes[v] >> 32) & (cap - 1);
                while (owner[h] >= 0 && keys[(size_t)owner[h]] != keys[v]) h = (h + 1) & (cap - 1);
                return h;
            };
            for (size_t v = 0; v < nv; ++v) {
                if (!mine(v)) continue;
                const size_t h = find(v);
                if (owner[h] < 0) {
                    owner[h] = (int)v;
                    least[h] = first[v];
                } else {
                    least[h] = std::min(least[h], first[v]);
                }
            }
            for (size_t v = 0; v < nv; ++v)
                if (mine(v)) rep[v] = least[find(v)];
        });
        auto corner_idx = [](size_t slot) { return (int)(4 * (slot / 3) + slot % 3); };
        run_slices(nc, threads, [&](size_t a, size_t e) {
            for (size_t slot = a; slot < e; ++slot) {
                const uint32_t v = mv.indices[slot];
                list[slot] = has_nan[v] ? corner_idx(slot) : corner_idx((size_t)rep[v]);
            }
        });
    }
    // 2. degenerate triangles (equal corner positions) go to the back; good ones keep their order
    std::vector<Tri> tris(nf);
    std::vector<int> order;  // list slot -> original face
    order.reserve(nf);
    for (int f = 0; f < nf; ++f) {
        V p0 = M.pos(list[3 * f]), p1 = M.pos(list[3 * f + 1]), p2 = M.pos(list[3 * f + 2]);
        if (!(veq(p0, p1) || veq(p0, p2) || veq(p1, p2))) order.push_back(f);
    }
    const int ngood = (int)order.size();
    for (int f = 0; f < nf; ++f) {
        V p0 = M.pos(list[3 * f]), p1 = M.pos(list[3 * f + 1]), p2 = M.pos(list[3 * f + 2]);
        if (veq(p0, p1) || veq(p0, p2) || veq(p1, p2)) order.push_back(f);
    }
    std::vector<int> tl((size_t)nf * 3);  // welded corners in list order
    for (int t = 0; t < nf; ++t) {
        for (int i = 0; i < 3; ++i) tl[3 * (size_t)t + i] = list[3 * (size_t)order[t] + i];
        tris[t].face = order[t];
        tris[t].flag = t >= ngood ? MARK_DEGENERATE : 0;
    }
    // 3. triangle derivative frames
    run_slices((size_t)ngood, ngood >= 65536 ? host_threads() : 1u, [&](size_t ta, size_t te) {
        for (int t = (int)ta; t < (int)te; ++t) {
            Tri& T = tris[t];
            T.flag |= GROUP_WITH_ANY;  // assumed bad until shown good
            const V v1 = M.pos(tl[3 * t]), v2 = M.pos(tl[3 * t + 1]), v3 = M.pos(tl[3 * t + 2]);
            const V t1 = M.tex(tl[3 * t]), t2 = M.tex(tl[3 * t + 1]), t3 = M.tex(tl[3 * t + 2]);
            const float t21x = t2.x - t1.x, t21y = t2.y - t1.y, t31x = t3.x - t1.x, t31y = t3.y - t1.y;
            const V d1 = vsub(v2, v1), d2 = vsub(v3, v1);
            const float area2 = t21x * t31y - t21y * t31x;
            const V os = vsub(vscale(t31y, d1), vscale(t21y, d2));
            const V ot = vadd(vscale(-t31x, d1), vscale(t21x, d2));
            if (area2 > 0) T.flag |= ORIENT_PRESERVING;
            if (nz(area2)) {
                const float abs_area = std::fabs(area2);
                const float len_os = vlen(os), len_ot = vlen(ot);
                const float s = (T.flag & ORIENT_PRESERVING) ? 1.0f : -1.0f;
                if (nz(len_os)) T.os = vscale(s / len_os, os);
                if (nz(len_ot)) T.ot = vscale(s / len_ot, ot);
                T.mag_s = len_os / abs_area;
                T.mag_t = len_ot / abs_area;
                if (nz(T.mag_s) && nz(T.mag_t)) T.flag &= ~GROUP_WITH_ANY;
            }
        }
    });
    // 4. neighbours: edges keyed (min, max, triangle), fully sorted
    struct Edge {
        int i0, i1, f;
    };
    std::vector<Edge> edges((size_t)ngood * 3);
    for (int t = 0; t < ngood; ++t)
        for (int i = 0; i < 3; ++i) {
            int a = tl[3 * t + i], b = tl[3 * t + (i < 2 ? i + 1 : 0)];
            edges[3 * (size_t)t + i] = {std::min(a, b), std::max(a, b), t};
        }
    std::sort(edges.begin(), edges.end(), [](const Edge& a, const Edge& b) {
        if (a.i0 != b.i0) return a.i0 < b.i0;
        if (a.i1 != b.i1) return a.i1 < b.i1;
        return a.f < b.f;
    });
    auto get_edge = [&](int t, int e0, int e1, int& o0, int& o1, int& num) {
        const int* ix = &tl[3 * (size_t)t];
        if (ix[0] == e0 || ix[0] == e1) {
            if (ix[1] == e0 || ix[1] == e1) {
                num = 0;
                o0 = ix[0];
                o1 = ix[1];
            } else {
                num = 2;
                o0 = ix[2];
                o1 = ix[0];
            }
        } else {
            num = 1;
            o0 = ix[1];
            o1 = ix[2];
        }
    };
    for (size_t i = 0; i < edges.size(); ++i) {
        const Edge& e = edges[i];
        int a0, a1, na;
        get_edge(e.f, e.i0, e.i1, a0, a1, na);
        if (tris[e.f].neighbors[na] != -1) continue;
        for (size_t j = i + 1; j < edges.size() && edges[j].i0 == e.i0 && edges[j].i1 == e.i1; ++j) {
            int b0, b1, nb;
            get_edge(edges[j].f, edges[j].i0, edges[j].i1, b1, b0, nb);  // flipped
            if (a0 == b0 && a1 == b1 && tris[edges[j].f].neighbors[nb] == -1) {
                tris[e.f].neighbors[na] = edges[j].f;
                tris[edges[j].f].neighbors[nb] = e.f;
                break;
            }
        }
    }
    // 5. groups around each welded vertex
    std::vector<Group> groups;
    groups.reserve((size_t)ngood);
    // depth-first assignment through neighbours of equal orientation
    std::vector<int> stack;
    auto assign = [&](int t0, int g) {
        // recursive in the method (left neighbour, then right); an explicit stack keeps that order
        stack.clear();
        stack.push_back(t0);
        while (!stack.empty()) {
            const int t = stack.back();
            stack.pop_back();
            Tri& T = tris[t];
            const int rep = groups[g].rep;
            int i = tl[3 * t] == rep ? 0 : tl[3 * t + 1] == rep ? 1 : tl[3 * t + 2] == rep ? 2 : -1;
            if (i < 0) continue;
            if (T.group[i] != -1) continue;  // in this group already, or in another
            if ((T.flag & GROUP_WITH_ANY) && T.group[0] == -1 && T.group[1] == -1 && T.group[2] == -1) {
                // the first group to take a group-with-anything triangle sets its orientation
                T.flag &= ~ORIENT_PRESERVING;
                T.flag |= groups[g].orient ? ORIENT_PRESERVING : 0;
            }
            if (((T.flag & ORIENT_PRESERVING) != 0) != groups[g].orient) continue;
            groups[g].faces.push_back(t);
            T.group[i] = g;
            const int nl = T.neighbors[i], nr = T.neighbors[i > 0 ? i - 1 : 2];
            // left is visited before right: push right first
            if (nr >= 0) stack.push_back(nr);
            if (nl >= 0) stack.push_back(nl);
        }
    };
    for (int t = 0; t < ngood; ++t)
        for (int i = 0; i < 3; ++i) {
            Tri& T = tris[t];
            if ((T.flag & GROUP_WITH_ANY) || T.group[i] != -1) continue;
            const int g = (int)groups.size();
            groups.push_back(Group{tl[3 * t + i], (T.flag & ORIENT_PRESERVING) != 0, {}});
            groups[g].faces.push_back(t);
            T.group[i] = g;
            const int nl = T.neighbors[i], nr = T.neighbors[i > 0 ? i - 1 : 2];
            if (nl >= 0) assign(nl, g);
            if (nr >= 0) assign(nr, g);
        }
    // 6. frames: within a group, each member's subgroup is the members whose
    //    projected derivatives are not exactly opposite to its own (cosine > -1,
    //    the 180-degree threshold) or that group with anything; equal subgroups
    //    share one frame, the angle-weighted mean over members in ascending order
    std::vector<Frame> frames((size_t)nf * 3);  // by original face * 3 + corner
    auto corner_of = [&](int t, int rep) { return tl[3 * t] == rep ? 0 : tl[3 * t + 1] == rep ? 1 : 2; };
    auto project = [&](V v, V n) {
        V p = vsub(v, vscale(vdot(n, v), n));
        return vnz(p) ? vnorm(p) : p;
    };
    auto eval_frame = [&](const std::vector<int>& members, int rep) {
        V sum_s{0, 0, 0}, sum_t{0, 0, 0};
        for (int t : members) {
            const Tri& T = tris[t];
            if (T.flag & GROUP_WITH_ANY) continue;
            const int i = corner_of(t, rep);
            const V n = M.nrm(tl[3 * t + i]);
            const V os = project(T.os, n), ot = project(T.ot, n);
            const int i2 = tl[3 * t + (i < 2 ? i + 1 : 0)], i1 = tl[3 * t + i], i0 = tl[3 * t + (i > 0 ? i - 1 : 2)];
            const V e1 = project(vsub(M.pos(i0), M.pos(i1)), n), e2 = project(vsub(M.pos(i2), M.pos(i1)), n);
            float c = vdot(e1, e2);
            c = c > 1 ? 1 : (c < -1 ? -1 : c);
            const float angle = std::acos(c);
            sum_s = vadd(sum_s, vscale(angle, os));
            sum_t = vadd(sum_t, vscale(angle, ot));
        }
        Frame fr;
        fr.os = vnz(sum_s) ? vnorm(sum_s) : sum_s;
        fr.ot = vnz(sum_t) ? vnorm(sum_t) : sum_t;
        return fr;
    };
    // scratch reused across groups: the members' derivatives projected once
    // into the group's tangent plane, and the distinct subgroups so far
    std::vector<V> proj_s, proj_t;
    std::vector<int> members;
    std::vector<std::vector<int>> subgroups;
    std::vector<Frame> sub_frames;
    for (const Group& G : groups) {
        const V n = M.nrm(G.rep);
        const size_t k = G.faces.size();
        proj_s.resize(k);
        proj_t.resize(k);
        for (size_t j = 0; j < k; ++j) {
            proj_s[j] = project(tris[G.faces[j]].os, n);
            proj_t[j] = project(tris[G.faces[j]].ot, n);
        }
        size_t nsub = 0;
        for (size_t a = 0; a < k; ++a) {
            const int f = G.faces[a];
            const Tri& F = tris[f];
            members.clear();
            for (size_t b = 0; b < k; ++b) {
                const int t = G.faces[b];
                const Tri& T = tris[t];
                const bool any = ((F.flag | T.flag) & GROUP_WITH_ANY) != 0;
                if (any || F.face == T.face ||
                    (vdot(proj_s[a], proj_s[b]) > -1.0f && vdot(proj_t[a], proj_t[b]) > -1.0f))
                    members.push_back(t);
            }
            std::sort(members.begin(), members.end());
            size_t l = 0;
            while (l < nsub && subgroups[l] != members) ++l;
            if (l == nsub) {
                if (nsub == subgroups.size()) {
                    subgroups.push_back(members);
                    sub_frames.push_back(eval_frame(members, G.rep));
                } else {
                    subgroups[nsub] = members;
                    sub_frames[nsub] = eval_frame(members, G.rep);
                }
                ++nsub;
            }
            Frame fr = sub_frames[l];
            fr.orient = G.orient;
            frames[3 * (size_t)F.face + corner_of(f, G.rep)] = fr;
        }
    }
    // 7. degenerate triangles copy the first good corner with the same welded vertex
    if (ngood < nf) {
        std::vector<int> first_good(4 * (size_t)nf, -1);  // welded vertex (a corner idx) -> list slot
        for (int j = 3 * ngood - 1; j >= 0; --j) first_good[(size_t)tl[j]] = j;
        for (int t = ngood; t < nf; ++t)
            for (int i = 0; i < 3; ++i) {
                const int src = first_good[(size_t)tl[3 * t + i]];
                if (src < 0) continue;
                frames[3 * (size_t)tris[t].face + i] = frames[3 * (size_t)tris[src / 3].face + src % 3];
            }
    }
    for (size_t k = 0; k < frames.size(); ++k) {
        out[4 * k + 0] = frames[k].os.x;
        out[4 * k + 1] = frames[k].os.y;
        out[4 * k + 2] = frames[k].os.z;
        out[4 * k + 3] = frames[k].orient ? 1.0f : -1.0f;
    }
    return true;
}

}  // namespace ptcore
