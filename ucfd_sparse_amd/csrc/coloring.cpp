// Greedy multi-colouring and colour-contiguous renumbering of a cell mesh.
//
// The reference's own colouring routine is absent from its tree; its only
// colouring code is the red-black helper examples/utils/pbutils.c:98-121.
// The routine here is first-fit in ascending cell id: cell i takes the
// smallest colour (1-based) not used by an already coloured face neighbour
// (a neighbour with a smaller id), on the symmetrised face graph; boundary
// self-faces are ignored.  On an unpermuted structured hex grid this is
// exactly the red-black colouring and icolor order of pbutils, and it never
// needs more than nface + 1 colours.  icolor lists the cells colour by
// colour, ascending id within a colour (the convention of
// examples/omp3d.c:173-175).
#include "ucfd_internal.h"
#include "ucfd_ext.h"

#include <algorithm>
#include <numeric>
#include <vector>

namespace {

// symmetrised adjacency (CSR) of the face graph, self faces dropped
void sym_graph(int N, int F, const int *nei, std::vector<int> &ptr, std::vector<int> &adj)
{
    std::vector<int> deg(N, 0);
    for (int f = 0; f < F; ++f)
        for (int i = 0; i < N; ++i) {
            const int nb = nei[(size_t)f*N + i];
            if (nb < 0 || nb >= N || nb == i) continue;
            deg[i]++;
            deg[nb]++;
        }
    ptr.assign(N + 1, 0);
    for (int i = 0; i < N; ++i) ptr[i + 1] = ptr[i] + deg[i];
    adj.assign(ptr[N], 0);
    std::vector<int> pos(ptr.begin(), ptr.end() - 1);
    for (int f = 0; f < F; ++f)
        for (int i = 0; i < N; ++i) {
            const int nb = nei[(size_t)f*N + i];
            if (nb < 0 || nb >= N || nb == i) continue;
            adj[pos[i]++] = nb;
            adj[pos[nb]++] = i;
        }
}

}  // namespace

namespace {

// first-fit in the given visiting order (nullptr: ascending id)
int greedy(int N, int nface, const int *nei_ele, const int *order, int *icolor, int *lcolor,
           int *color_off)
{
    std::vector<int> ptr, adj;
    sym_graph(N, nface, nei_ele, ptr, adj);
    std::vector<int> mark(2*nface + 8, -1);
    std::vector<char> done(N, 0);
    int ncol = 0;
    for (int k = 0; k < N; ++k) {
        const int i = order ? order[k] : k;
        for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
            const int nb = adj[e];
            if (done[nb]) {
                const int c = lcolor[nb];
                if (c >= (int)mark.size()) mark.resize(2*c + 2, -1);
                mark[c] = i;
            }
        }
        int c = 1;
        while (c < (int)mark.size() && mark[c] == i) ++c;
        if (c >= (int)mark.size()) mark.resize(2*c + 2, -1);
        lcolor[i] = c;
        done[i] = 1;
        ncol = std::max(ncol, c);
    }
    std::vector<int> off(ncol + 2, 0);
    for (int i = 0; i < N; ++i) off[lcolor[i] + 1]++;
    for (int c = 1; c <= ncol; ++c) off[c + 1] += off[c];
    if (icolor) {
        std::vector<int> pos(off.begin(), off.end() - 1);
        for (int i = 0; i < N; ++i) icolor[pos[lcolor[i]]++] = i;
    }
    if (color_off)
        for (int c = 0; c <= ncol; ++c) color_off[c] = off[c + 1];
    return ncol;
}

}  // namespace

// Returns the number of colours (>= 1), or a negative value on bad input.
UCFD_API int ucfd_greedy_coloring(int neles, int nface, const int *nei_ele, int *icolor,
                                  int *lcolor, int *color_off)
{
    if (neles < 0 || nface < 1 || !nei_ele || !lcolor) return -1;
    if (neles == 0) {
        if (color_off) color_off[0] = 0;
        return 0;
    }
    return greedy(neles, nface, nei_ele, nullptr, icolor, lcolor, color_off);
}

// First-fit visiting the cells in order[0..neles) (a permutation), e.g. along
// a space-filling curve: spatially coherent visiting packs most cells into
// the first two colours, and the remaining sparse colours are small.
UCFD_API int ucfd_greedy_coloring_ordered(int neles, int nface, const int *nei_ele,
                                          const int *order, int *icolor, int *lcolor,
                                          int *color_off)
{
    if (neles < 0 || nface < 1 || !nei_ele || !lcolor || !order) return -1;
    if (neles == 0) {
        if (color_off) color_off[0] = 0;
        return 0;
    }
    std::vector<char> seen(neles, 0);
    for (int k = 0; k < neles; ++k) {
        if (order[k] < 0 || order[k] >= neles || seen[order[k]]) return -1;
        seen[order[k]] = 1;
    }
    return greedy(neles, nface, nei_ele, order, icolor, lcolor, color_off);
}

// Breadth-first visiting order of the face graph: components are entered at
// their first cell in seeds[0..neles) (ascending id when seeds is NULL) and
// each cell's neighbours are queued in face order.  First-fit colouring in
// this order gives every cell of BFS layer k a neighbour in layer k-1, so on
// a bipartite face graph (structured hex grids, Kuhn/Freudenthal tet splits)
// it returns the 2-colouring; on other meshes it is one more greedy order.
// Returns the number of connected components, or -1 on bad input.
UCFD_API int ucfd_bfs_order(int neles, int nface, const int *nei_ele, const int *seeds,
                            int *order)
{
    if (neles < 0 || nface < 1 || !nei_ele || !order) return -1;
    if (seeds) {
        std::vector<char> seen(neles, 0);
        for (int k = 0; k < neles; ++k) {
            if (seeds[k] < 0 || seeds[k] >= neles || seen[seeds[k]]) return -1;
            seen[seeds[k]] = 1;
        }
    }
    std::vector<int> ptr, adj;
    sym_graph(neles, nface, nei_ele, ptr, adj);
    std::vector<char> vis(neles, 0);
    int head = 0, tail = 0, ncomp = 0;
    for (int k = 0; k < neles; ++k) {
        const int s = seeds ? seeds[k] : k;
        if (vis[s]) continue;
        ++ncomp;
        vis[s] = 1;
        order[tail++] = s;
        while (head < tail) {
            const int i = order[head++];
            for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
                const int nb = adj[e];
                if (!vis[nb]) {
                    vis[nb] = 1;
                    order[tail++] = nb;
                }
            }
        }
    }
    return ncomp;
}

// Balanced colouring: cells in ascending id take, among the colours 1..C not
// used by an already coloured face neighbour, the one holding the fewest
// cells so far (ties: the smallest colour).  C = max(ncolors, maximum degree
// + 1), so a free colour always exists.  Colours come out of nearly equal
// size: no small trailing colour, and every colour dense enough that the
// sweeps' neighbour gathers share cache lines.  Returns C.
UCFD_API int ucfd_balanced_coloring(int neles, int nface, const int *nei_ele, int ncolors,
                                    int *icolor, int *lcolor, int *color_off)
{
    if (neles < 0 || nface < 1 || !nei_ele || !lcolor || ncolors < 0) return -1;
    if (neles == 0) {
        if (color_off) color_off[0] = 0;
        return 0;
    }
    const int N = neles;
    std::vector<int> ptr, adj;
    sym_graph(N, nface, nei_ele, ptr, adj);
    int maxdeg = 0;     // distinct neighbours (each face pair is listed from both sides)
    for (int i = 0; i < N; ++i) {
        std::sort(adj.begin() + ptr[i], adj.begin() + ptr[i + 1]);
        const int d = (int)(std::unique(adj.begin() + ptr[i], adj.begin() + ptr[i + 1]) -
                            (adj.begin() + ptr[i]));
        maxdeg = std::max(maxdeg, d);
    }
    const int C = std::max(ncolors, maxdeg + 1);
    std::vector<long long> count(C + 1, 0);
    std::vector<int> mark(C + 1, -1);
    for (int i = 0; i < N; ++i) {
        for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
            const int nb = adj[e];
            if (nb < i) mark[lcolor[nb]] = i;
        }
        int best = 0;
        for (int c = 1; c <= C; ++c)
            if (mark[c] != i && (best == 0 || count[c] < count[best])) best = c;
        lcolor[i] = best;
        count[best]++;
    }
    std::vector<int> off(C + 2, 0);
    for (int i = 0; i < N; ++i) off[lcolor[i] + 1]++;
    for (int c = 1; c <= C; ++c) off[c + 1] += off[c];
    if (icolor) {
        std::vector<int> pos(off.begin(), off.end() - 1);
        for (int i = 0; i < N; ++i) icolor[pos[lcolor[i]]++] = i;
    }
    if (color_off)
        for (int c = 0; c <= C; ++c) color_off[c] = off[c + 1];
    return C;
}

// Removes the highest colour when possible: each of its cells takes a colour
// no neighbour uses, or one freed by swapping a two-colour Kempe chain
// (colours a, b) that starts at its only a-coloured neighbour and holds none
// of its b-coloured neighbours; chains longer than max_chain cells are not
// tried.  A cell left over keeps the top colour.  lcolor is updated in
// place, icolor / color_off rebuilt.  Returns the new number of colours.
UCFD_API int ucfd_reduce_colors(int neles, int nface, const int *nei_ele, int max_chain,
                                int *icolor, int *lcolor, int *color_off)
{
    if (neles < 0 || nface < 1 || !nei_ele || !lcolor || max_chain < 1) return -1;
    const int N = neles;
    if (N == 0) return 0;
    std::vector<int> ptr0, adj0, ptr(N + 1, 0), adj;
    sym_graph(N, nface, nei_ele, ptr0, adj0);
    for (int i = 0; i < N; ++i) {          // distinct neighbours only
        std::sort(adj0.begin() + ptr0[i], adj0.begin() + ptr0[i + 1]);
        const auto end = std::unique(adj0.begin() + ptr0[i], adj0.begin() + ptr0[i + 1]);
        adj.insert(adj.end(), adj0.begin() + ptr0[i], end);
        ptr[i + 1] = (int)adj.size();
    }
    int top = 0;
    for (int i = 0; i < N; ++i) top = std::max(top, lcolor[i]);
    if (top <= 1) return top;
    std::vector<int> stamp(N, -1), comp;
    int tag = 0;
    for (int v = 0; v < N; ++v) {
        if (lcolor[v] != top) continue;
        std::vector<int> cnt(top + 1, 0);
        for (int e = ptr[v]; e < ptr[v + 1]; ++e)
            if (adj[e] != v) cnt[lcolor[adj[e]]]++;
        int c = 1;
        while (c < top && cnt[c]) ++c;
        if (c < top) {
            lcolor[v] = c;
            continue;
        }
        bool done = false;
        for (int a = top - 1; a >= 1 && !done; --a) {
            int ua = -1;
            for (int e = ptr[v]; e < ptr[v + 1]; ++e)
                if (adj[e] != v && lcolor[adj[e]] == a) ua = adj[e];
            if (cnt[a] != 1 || ua < 0) continue;
            for (int b = top - 1; b >= 1 && !done; --b) {
                if (b == a) continue;
                ++tag;
                comp.assign(1, ua);
                stamp[ua] = tag;
                bool ok = true;
                for (size_t q = 0; q < comp.size() && ok; ++q) {
                    const int x = comp[q];
                    for (int e = ptr[x]; e < ptr[x + 1]; ++e) {
                        const int y = adj[e];
                        if (y == v || stamp[y] == tag || (lcolor[y] != a && lcolor[y] != b)) continue;
                        stamp[y] = tag;
                        comp.push_back(y);
                        if ((int)comp.size() > max_chain) { ok = false; break; }
                    }
                }
                if (!ok) continue;
                for (int e = ptr[v]; e < ptr[v + 1] && ok; ++e)
                    if (adj[e] != v && lcolor[adj[e]] == b && stamp[adj[e]] == tag) ok = false;
                if (!ok) continue;
                for (int x : comp) lcolor[x] = lcolor[x] == a ? b : a;
                lcolor[v] = a;
                done = true;
            }
        }
    }
    int ncol = 0;
    for (int i = 0; i < N; ++i) ncol = std::max(ncol, lcolor[i]);
    std::vector<int> off(ncol + 2, 0);
    for (int i = 0; i < N; ++i) off[lcolor[i] + 1]++;
    for (int c = 1; c <= ncol; ++c) off[c + 1] += off[c];
    if (icolor) {
        std::vector<int> pos(off.begin(), off.end() - 1);
        for (int i = 0; i < N; ++i) icolor[pos[lcolor[i]]++] = i;
    }
    if (color_off)
        for (int c = 0; c <= ncol; ++c) color_off[c] = off[c + 1];
    return ncol;
}

// Number of face pairs (i, nb != i) with lcolor[i] == lcolor[nb]; 0 means a
// valid colouring.  Out-of-range neighbours count as conflicts.
UCFD_API long long ucfd_coloring_conflicts(int neles, int nface, const int *nei_ele,
                                           const int *lcolor)
{
    long long bad = 0;
    const int N = neles;
    for (int f = 0; f < nface; ++f)
        for (int i = 0; i < N; ++i) {
            const int nb = nei_ele[(size_t)f*N + i];
            if (nb < 0 || nb >= N) { ++bad; continue; }
            if (nb != i && lcolor[nb] == lcolor[i]) ++bad;
        }
    return bad;
}

// Colour-contiguous renumbering: perm[old] = new position, with cells sorted
// by (lcolor, key, old id).  key may be NULL (then ascending id within a
// colour, i.e. new order = icolor).  color_off[c] .. color_off[c+1] is the
// new range of colour c+1 (ncolors+1 entries).  Returns ncolors.
UCFD_API int ucfd_color_reorder(int neles, const int *lcolor, const unsigned long long *key,
                                int *perm, int *color_off)
{
    if (neles < 0 || !lcolor || !perm) return -1;
    const int N = neles;
    std::vector<int> idx(N);
    std::iota(idx.begin(), idx.end(), 0);
    int ncol = 0;
    for (int i = 0; i < N; ++i) ncol = std::max(ncol, lcolor[i]);
    if (key)
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
            if (lcolor[a] != lcolor[b]) return lcolor[a] < lcolor[b];
            return key[a] < key[b];
        });
    else
        std::stable_sort(idx.begin(), idx.end(),
                         [&](int a, int b) { return lcolor[a] < lcolor[b]; });
    for (int q = 0; q < N; ++q) perm[idx[q]] = q;
    if (color_off) {
        std::vector<int> off(ncol + 2, 0);
        for (int i = 0; i < N; ++i) off[lcolor[i] + 1]++;
        for (int c = 1; c <= ncol; ++c) off[c + 1] += off[c];
        for (int c = 0; c <= ncol; ++c) color_off[c] = off[c + 1];
    }
    return ncol;
}

// out[r*N + perm[i]] = in[r*N + i] for rows r < nrows (element size esize).
UCFD_API int ucfd_permute_rows(int neles, int nrows, int esize, const int *perm, const void *in,
                               void *out)
{
    if (neles < 0 || nrows < 0 || (esize != 4 && esize != 8) || !perm || !in || !out) return 1;
    const size_t N = (size_t)neles;
    for (int r = 0; r < nrows; ++r) {
        if (esize == 8) {
            const double *a = (const double *)in + r*N;
            double *b = (double *)out + r*N;
            for (size_t i = 0; i < N; ++i) b[perm[i]] = a[i];
        } else {
            const int *a = (const int *)in + r*N;
            int *b = (int *)out + r*N;
            for (size_t i = 0; i < N; ++i) b[perm[i]] = a[i];
        }
    }
    return 0;
}

// Renumbered neighbour table: out[f*N + perm[i]] = perm[nei[f*N + i]].
UCFD_API int ucfd_permute_nei(int neles, int nface, const int *perm, const int *nei, int *out)
{
    if (neles < 0 || nface < 1 || !perm || !nei || !out) return 1;
    const size_t N = (size_t)neles;
    for (int f = 0; f < nface; ++f)
        for (size_t i = 0; i < N; ++i) {
            const int nb = nei[f*N + i];
            out[f*N + perm[i]] = (nb >= 0 && (size_t)nb < N) ? perm[nb] : nb;
        }
    return 0;
}
