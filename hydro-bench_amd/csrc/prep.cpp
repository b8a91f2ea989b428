// prep.cpp -- host-side mesh preparation (native C++, no device code).
#include "prep.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace hydro {

namespace {

struct Csr {
    std::vector<int32_t> start, adj;
};

// Undirected cell adjacency from the slot neighbour array.
Csr cell_graph(int C, const int32_t *NAC) {
    Csr g;
    g.start.assign(C + 1, 0);
    for (int i = 0; i < C; i++)
        for (int j = 0; j < 4; j++) {
            int n = NAC[4 * i + j] - 1;
            if (n >= 0 && n < C && n != i) g.start[i + 1]++;
        }
    for (int i = 0; i < C; i++) g.start[i + 1] += g.start[i];
    g.adj.resize(g.start[C]);
    std::vector<int32_t> fill(g.start.begin(), g.start.end() - 1);
    for (int i = 0; i < C; i++)
        for (int j = 0; j < 4; j++) {
            int n = NAC[4 * i + j] - 1;
            if (n >= 0 && n < C && n != i) g.adj[fill[i]++] = n;
        }
    return g;
}

// BFS levels from `root` restricted to unvisited-in-component cells; returns
// the last cell reached and the eccentricity.
std::pair<int, int> bfs_far(const Csr &g, int root, std::vector<int32_t> &level) {
    std::vector<int32_t> q;
    q.push_back(root);
    level[root] = 0;
    size_t h = 0;
    int last = root;
    while (h < q.size()) {
        int v = q[h++];
        last = v;
        for (int k = g.start[v]; k < g.start[v + 1]; k++) {
            int w = g.adj[k];
            if (level[w] < 0) {
                level[w] = level[v] + 1;
                q.push_back(w);
            }
        }
    }
    int ecc = level[last];
    for (int v : q) level[v] = -1;
    return {last, ecc};
}

}  // namespace

std::vector<int32_t> rcm_order(int C, const int32_t *NAC) {
    Csr g = cell_graph(C, NAC);
    std::vector<int32_t> deg(C);
    for (int i = 0; i < C; i++) deg[i] = g.start[i + 1] - g.start[i];
    std::vector<int32_t> order;
    order.reserve(C);
    std::vector<char> seen(C, 0);
    std::vector<int32_t> level(C, -1);
    std::vector<int32_t> nb;
    for (int s0 = 0; s0 < C; s0++) {
        if (seen[s0]) continue;
        // pseudo-peripheral start: repeat BFS to the farthest cell while the
        // eccentricity grows (George-Liu)
        int root = s0, ecc = -1;
        for (int it = 0; it < 8; it++) {
            auto far = bfs_far(g, root, level);
            if (far.second <= ecc) break;
            ecc = far.second;
            root = far.first;
        }
        size_t head = order.size();
        order.push_back(root);
        seen[root] = 1;
        while (head < order.size()) {
            int v = order[head++];
            nb.clear();
            for (int k = g.start[v]; k < g.start[v + 1]; k++) {
                int w = g.adj[k];
                if (!seen[w]) {
                    seen[w] = 1;
                    nb.push_back(w);
                }
            }
            std::sort(nb.begin(), nb.end(), [&](int a, int b) {
                return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
            });
            order.insert(order.end(), nb.begin(), nb.end());
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

std::vector<int32_t> tile_order(int C, const int32_t *NAC, int B, const std::vector<int32_t> &along) {
    Csr g = cell_graph(C, NAC);
    std::vector<int32_t> rank(C);
    for (int i = 0; i < C; i++) rank[along[i]] = i;
    std::vector<char> taken(C, 0);
    std::vector<int32_t> order;
    order.reserve(C);
    std::vector<int32_t> queue;
    size_t next = 0;   // first position of `along` that may still be free
    int hint = -1;     // seed proposed by the previous tile's frontier
    auto global_seed = [&]() {
        while (next < (size_t)C && taken[along[next]]) next++;
        return next < (size_t)C ? along[next] : -1;
    };
    while ((int)order.size() < C) {
        const size_t t0 = order.size();
        const size_t want = std::min<size_t>((size_t)B, (size_t)C - t0);
        int seed = (hint >= 0 && !taken[hint]) ? hint : global_seed();
        while (order.size() - t0 < want) {
            if (seed < 0) seed = global_seed();
            queue.clear();
            queue.push_back(seed);
            taken[seed] = 1;
            order.push_back(seed);
            seed = -1;
            for (size_t h = 0; h < queue.size() && order.size() - t0 < want; h++) {
                const int v = queue[h];
                for (int k = g.start[v]; k < g.start[v + 1] && order.size() - t0 < want; k++) {
                    const int w = g.adj[k];
                    if (taken[w]) continue;
                    taken[w] = 1;
                    order.push_back(w);
                    queue.push_back(w);
                }
            }
        }
        // next seed: the free neighbour of this tile earliest in `along`
        hint = -1;
        for (size_t i = t0; i < order.size(); i++) {
            const int v = order[i];
            for (int k = g.start[v]; k < g.start[v + 1]; k++) {
                const int w = g.adj[k];
                if (!taken[w] && (hint < 0 || rank[w] < rank[hint])) hint = w;
            }
        }
    }
    return order;
}

namespace {
// Hilbert index of (x, y) in a 2^order grid (iterative rotate-and-flip).
uint64_t hilbert_d(uint32_t n, uint32_t x, uint32_t y) {
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        uint32_t rx = (x & s) > 0, ry = (y & s) > 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}
}  // namespace

std::vector<int32_t> hilbert_order(int C, const double *xc, const double *yc) {
    std::vector<int32_t> perm(C);
    std::iota(perm.begin(), perm.end(), 0);
    if (C == 0) return perm;
    double x0 = xc[0], x1 = xc[0], y0 = yc[0], y1 = yc[0];
    for (int i = 1; i < C; i++) {
        x0 = std::min(x0, xc[i]); x1 = std::max(x1, xc[i]);
        y0 = std::min(y0, yc[i]); y1 = std::max(y1, yc[i]);
    }
    const uint32_t n = 1u << 16;
    const double span = std::max(x1 - x0, y1 - y0);
    const double sc = span > 0 ? (n - 1) / span : 0.0;
    std::vector<uint64_t> key(C);
    for (int i = 0; i < C; i++) {
        uint32_t xi = (uint32_t)std::lround((xc[i] - x0) * sc);
        uint32_t yi = (uint32_t)std::lround((yc[i] - y0) * sc);
        key[i] = hilbert_d(n, xi, yi);
    }
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    return perm;
}

template <class R>
void precalculate(int CELL, int NOD, const int32_t *NAP, R *XP, R *YP, const R *ZBC, R *Z, const R *FNC0,
                  R HM1, R *SIDE, R *SINF, R *COSF, R *SLCOS, R *SLSIN, R *AREA, R *ZB1, R *H, R *FNC) {
    if (NOD > 0) {
        R xmin = *std::min_element(XP, XP + NOD);
        R ymin = *std::min_element(YP, YP + NOD);
        for (int i = 0; i < NOD; i++) {
            XP[i] -= xmin;
            YP[i] -= ymin;
        }
    }
    std::fill(SIDE, SIDE + 4 * (size_t)CELL, (R)0);
    std::fill(SINF, SINF + 4 * (size_t)CELL, (R)0);
    std::fill(COSF, COSF + 4 * (size_t)CELL, (R)0);
    std::fill(SLCOS, SLCOS + 4 * (size_t)CELL, (R)0);
    std::fill(SLSIN, SLSIN + 4 * (size_t)CELL, (R)0);
    std::fill(AREA, AREA + CELL, (R)0);
    for (int i = 0; i < CELL; i++) {
        const int32_t *nw = NAP + 4 * (size_t)i;
        if (nw[0] == 0) continue;
        const int nv = (nw[3] == 0 || nw[3] == nw[0]) ? 3 : 4;
        const R X1 = XP[nw[0] - 1], X2 = XP[nw[1] - 1], X3 = XP[nw[2] - 1];
        const R Y1 = YP[nw[0] - 1], Y2 = YP[nw[1] - 1], Y3 = YP[nw[2] - 1];
        // area of a triangle fan from node 1, in the reference's operand order
        AREA[i] = std::fabs(((Y3 - Y1) * (X2 - X1) - (X3 - X1) * (Y2 - Y1)) / 2.0);
        if (nv == 4) {
            const R X4 = XP[nw[3] - 1], Y4 = YP[nw[3] - 1];
            AREA[i] += std::fabs(((Y4 - Y1) * (X3 - X1) - (X4 - X1) * (Y3 - Y1)) / 2.0);
        }
        for (int j = 0; j < nv; j++) {
            const int a = nw[j] - 1, b = nw[(j + 1) % nv] - 1;
            const R DX = XP[a] - XP[b];
            const R DY = YP[b] - YP[a];
            const size_t s = 4 * (size_t)i + j;
            SIDE[s] = std::sqrt(DX * DX + DY * DY);
            if (SIDE[s] > 0.0) {
                SINF[s] = DX / SIDE[s];
                COSF[s] = DY / SIDE[s];
            }
        }
    }
    for (int i = 0; i < CELL; i++) ZB1[i] = ZBC[i] + HM1;
    for (int i = 0; i < CELL; i++) {
        H[i] = 0;
        if (NAP[4 * (size_t)i] == 0) continue;
        if (Z[i] <= ZBC[i]) {
            H[i] = HM1;
            Z[i] = ZB1[i];
        } else {
            H[i] = Z[i] - ZBC[i];
        }
    }
    for (int i = 0; i < CELL; i++) {
        FNC[i] = 9.81 * FNC0[i] * FNC0[i];
        const int32_t *nw = NAP + 4 * (size_t)i;
        const int nv = (nw[0] == 0) ? 0 : (nw[3] == 0 || nw[3] == nw[0]) ? 3 : 4;
        for (int j = 0; j < nv; j++) {
            const size_t s = 4 * (size_t)i + j;
            SLCOS[s] = SIDE[s] * COSF[s];
            SLSIN[s] = SIDE[s] * SINF[s];
        }
    }
}

template void precalculate<double>(int, int, const int32_t *, double *, double *, const double *, double *,
                                   const double *, double, double *, double *, double *, double *,
                                   double *, double *, double *, double *, double *);
template void precalculate<float>(int, int, const int32_t *, float *, float *, const float *, float *,
                                  const float *, float, float *, float *, float *, float *, float *,
                                  float *, float *, float *, float *);

// Within each tile of B consecutive positions of `perm`, the cells that
// share an edge with another tile (a neighbour there, or listed as a
// neighbour by a cell there) first, the others after them, each group in its
// previous order.  Tile membership does not change.
void edge_cells_first(int C, const int32_t *NAC, int B, std::vector<int32_t> &perm, const uint8_t *shallow) {
    if (B <= 0) return;
    std::vector<int32_t> inv(C);
    for (int p = 0; p < C; p++) inv[perm[p]] = p;
    std::vector<uint8_t> out(C, 0);   // by original id
    for (int o = 0; o < C; o++)
        for (int j = 0; j < 4; j++) {
            const int n = NAC[4 * (size_t)o + j] - 1;
            if (n >= 0 && inv[n] / B != inv[o] / B) out[o] = out[n] = 1;
        }
    std::vector<int32_t> edge, inner;
    for (int p0 = 0; p0 < C; p0 += B) {
        const int pend = std::min(C, p0 + B);
        edge.clear();
        inner.clear();
        for (int p = p0; p < pend; p++) (out[perm[p]] ? edge : inner).push_back(perm[p]);
        if (shallow) {
            auto deep_first = [&](int32_t a, int32_t b) { return shallow[a] < shallow[b]; };
            std::stable_sort(edge.begin(), edge.end(), deep_first);
            std::stable_sort(inner.begin(), inner.end(), deep_first);
        }
        std::copy(edge.begin(), edge.end(), perm.begin() + p0);
        std::copy(inner.begin(), inner.end(), perm.begin() + p0 + (int)edge.size());
    }
}

}  // namespace hydro
