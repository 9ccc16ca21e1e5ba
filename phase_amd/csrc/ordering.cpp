// Host-side ordering of a sparsity pattern for the device solver.
//
// The reference hands its CSR (rows in cell order, src/Math/CrsEquation.cpp:9-14)
// to Ifpack2's RILUK inside an additive Schwarz preconditioner
// (src/Math/TrilinosBelosSparseMatrixSolver.cpp:69-82), which factors rows in
// their natural order -- a strictly sequential recurrence.  For the GPU the
// rows are renumbered once per pattern:
//
//   1. reverse Cuthill-McKee over the symmetrised graph (locality);
//   2. rows grouped into blocks of <= 64 graph-connected rows, grown
//      breadth-first from seeds taken in RCM order (one wavefront per block);
//   3. blocks coloured greedily so that blocks of one colour share no
//      coupling;
//   4. solver order = blocks by (colour, creation index), rows inside a block
//      in RCM order.
//
// ILU(0) is then computed exactly (no dropped couplings) in that order.  A
// row depends only on earlier-coloured neighbour blocks and on earlier rows of
// its own block, so one wavefront per block can run the in-block recurrence by
// levels while blocks wait for their (few) predecessor blocks.
#include <algorithm>
#include <climits>
#include <cmath>
#include <random>
#include <cstring>
#include <numeric>
#include <queue>
#include <tuple>

#include "phase_internal.h"

namespace phase {

namespace {

struct Graph {
  std::vector<int> ptr, adj;
  int deg(int i) const { return ptr[i + 1] - ptr[i]; }
};

Graph symmetrise(int n, const std::vector<int> &rp, const std::vector<int> &ci) {
  std::vector<int> cnt(n, 0);
  for (int i = 0; i < n; ++i)
    for (int j = rp[i]; j < rp[i + 1]; ++j)
      if (ci[j] != i) { ++cnt[i]; ++cnt[ci[j]]; }
  Graph g;
  g.ptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) g.ptr[i + 1] = g.ptr[i] + cnt[i];
  std::vector<int> fill(g.ptr.begin(), g.ptr.end() - 1);
  g.adj.resize(g.ptr[n]);
  for (int i = 0; i < n; ++i)
    for (int j = rp[i]; j < rp[i + 1]; ++j)
      if (ci[j] != i) { g.adj[fill[i]++] = ci[j]; g.adj[fill[ci[j]]++] = i; }
  // sort + unique each row, then compact
  std::vector<int> ptr2(n + 1, 0);
  int w = 0;
  for (int i = 0; i < n; ++i) {
    int b = g.ptr[i], e = g.ptr[i + 1];
    std::sort(g.adj.begin() + b, g.adj.begin() + e);
    int last = -1;
    ptr2[i] = w;
    for (int k = b; k < e; ++k)
      if (g.adj[k] != last) { last = g.adj[k]; g.adj[w++] = last; }
  }
  ptr2[n] = w;
  g.adj.resize(w);
  g.ptr = ptr2;
  return g;
}

// BFS level structure from root; returns the last level's nodes and depth.
int bfsDepth(const Graph &g, int root, std::vector<int> &mark, int stamp, std::vector<int> &last) {
  std::vector<int> cur{root}, nxt;
  mark[root] = stamp;
  int depth = 0;
  while (true) {
    nxt.clear();
    for (int u : cur)
      for (int k = g.ptr[u]; k < g.ptr[u + 1]; ++k) {
        int w = g.adj[k];
        if (mark[w] != stamp) { mark[w] = stamp; nxt.push_back(w); }
      }
    if (nxt.empty()) break;
    cur.swap(nxt);
    ++depth;
  }
  last = cur;
  return depth;
}

std::vector<int> rcm(const Graph &g, int n) {
  std::vector<int> order;
  order.reserve(n);
  std::vector<char> visited(n, 0);
  std::vector<int> mark(n, -1), last;
  int stamp = 0;
  std::vector<int> byDeg(n);
  std::iota(byDeg.begin(), byDeg.end(), 0);
  std::stable_sort(byDeg.begin(), byDeg.end(), [&](int a, int b) { return g.deg(a) < g.deg(b); });
  std::vector<int> nb;
  for (int cand : byDeg) {
    if (visited[cand]) continue;
    // pseudo-peripheral root (George & Liu): walk to a far, low-degree node
    int root = cand;
    int depth = bfsDepth(g, root, mark, stamp++, last);
    for (int rep = 0; rep < 4; ++rep) {
      int best = last[0];
      for (int v : last)
        if (g.deg(v) < g.deg(best)) best = v;
      int d2 = bfsDepth(g, best, mark, stamp++, last);
      if (d2 <= depth) break;
      depth = d2;
      root = best;
    }
    size_t head = order.size();
    order.push_back(root);
    visited[root] = 1;
    while (head < order.size()) {
      int u = order[head++];
      nb.clear();
      for (int k = g.ptr[u]; k < g.ptr[u + 1]; ++k)
        if (!visited[g.adj[k]]) { visited[g.adj[k]] = 1; nb.push_back(g.adj[k]); }
      std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return g.deg(a) < g.deg(b); });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

}  // namespace

// Remove the top colour k-1 by Kempe-chain interchanges (the recolouring step
// of Kempe's 1879 argument): a block b of colour k-1 takes colour a < k-1 once
// the (a, c)-chains through its a-coloured neighbours are swapped, provided no
// such chain also reaches a c-coloured neighbour of b.  Returns the number of
// blocks left with colour k-1 (stopping once more than maxLeft are); `col` is a
// proper colouring either way.
static int kempeRecolour(const std::vector<std::vector<int>> &adj, int k, std::vector<int> &col, int maxLeft) {
  const int nb = (int)adj.size();
  std::vector<int> mark(nb, -1), chain, stack;
  int stamp = 0;
  int left = 0;
  for (int b = 0; b < nb; ++b) {
    if (col[b] != k - 1) continue;
    bool done = false;
    for (int a = 0; a < k - 1 && !done; ++a) {
      bool freeA = true;
      for (int w : adj[b]) freeA = freeA && col[w] != a;
      if (freeA) {
        col[b] = a;
        done = true;
        break;
      }
      for (int c = 0; c < k - 1 && !done; ++c) {
        if (c == a) continue;
        // union of the (a, c)-chains through b's a-coloured neighbours
        ++stamp;
        chain.clear();
        for (int w : adj[b]) {
          if (col[w] != a || mark[w] == stamp) continue;
          mark[w] = stamp;
          stack.assign(1, w);
          while (!stack.empty()) {
            const int u = stack.back();
            stack.pop_back();
            chain.push_back(u);
            for (int x : adj[u])
              if (x != b && mark[x] != stamp && (col[x] == a || col[x] == c)) {
                mark[x] = stamp;
                stack.push_back(x);
              }
          }
        }
        bool ok = true;
        for (int w : adj[b]) ok = ok && !(col[w] == c && mark[w] == stamp);
        if (!ok) continue;
        for (int u : chain) col[u] = col[u] == a ? c : a;
        col[b] = a;
        done = true;
      }
    }
    left += done ? 0 : 1;
    if (left > maxLeft) break;  // hopeless: stop early
  }
  return left;
}

// Tabucol (Hertz & de Werra 1987): look for a proper colouring with colours
// 0..k-1 starting from `col` (entries >= k first move to their least-conflicting
// colour).  A move recolours one conflicting vertex to the colour that lowers
// the conflict count most; the reverse move is tabu for 0.6 * conflicts +
// U{0..9} iterations unless it reaches zero conflicts.  Deterministic (fixed
// seed).  Returns true with the colouring in `col` on success within maxIters.
static bool tabuRecolour(const std::vector<std::vector<int>> &adj, int k, std::vector<int> &col, long maxIters) {
  const int nb = (int)adj.size();
  std::vector<int> gam((size_t)nb * k, 0);  // gam[v*k+c]: neighbours of v with colour c
  for (int v = 0; v < nb; ++v)
    if (col[v] >= k) col[v] = -1;
  for (int v = 0; v < nb; ++v) {
    if (col[v] >= 0) continue;
    std::vector<int> cnt(k, 0);
    for (int w : adj[v])
      if (col[w] >= 0) ++cnt[col[w]];
    col[v] = (int)(std::min_element(cnt.begin(), cnt.end()) - cnt.begin());
  }
  long f = 0;
  for (int v = 0; v < nb; ++v)
    for (int w : adj[v]) {
      ++gam[(size_t)v * k + col[w]];
      if (w > v && col[w] == col[v]) ++f;
    }
  std::vector<int> conf, pos(nb, -1);  // conflicting vertices, as an indexed set
  auto upd = [&](int v) {
    const bool in = gam[(size_t)v * k + col[v]] > 0;
    if (in && pos[v] < 0) {
      pos[v] = (int)conf.size();
      conf.push_back(v);
    } else if (!in && pos[v] >= 0) {
      const int last = conf.back();
      conf[pos[v]] = last;
      pos[last] = pos[v];
      conf.pop_back();
      pos[v] = -1;
    }
  };
  for (int v = 0; v < nb; ++v) upd(v);
  std::vector<long> tabu((size_t)nb * k, 0);
  std::mt19937 rng(12345);
  for (long it = 0; it < maxIters && f > 0; ++it) {
    int bv = -1, bc = -1;
    long bd = LONG_MAX;
    for (int v : conf) {
      const int cv = col[v];
      const int own = gam[(size_t)v * k + cv];
      for (int c = 0; c < k; ++c) {
        if (c == cv) continue;
        const long d = gam[(size_t)v * k + c] - own;
        if (tabu[(size_t)v * k + c] > it && f + d > 0) continue;
        if (d < bd) {
          bd = d;
          bv = v;
          bc = c;
        }
      }
    }
    if (bv < 0) continue;  // every move tabu: wait for tenures to expire
    const int old = col[bv];
    col[bv] = bc;
    f += bd;
    tabu[(size_t)bv * k + old] = it + (long)(0.6 * f) + (long)(rng() % 10);
    for (int w : adj[bv]) {
      --gam[(size_t)w * k + old];
      ++gam[(size_t)w * k + bc];
      upd(w);
    }
    upd(bv);
  }
  return f == 0;
}

// Brick aggregates for a 2-D mesh whose cell centroids are known: rows fall
// into rectangular bricks of about `target` rows laid as a running-bond wall
// (every other row of bricks shifted by half a brick).  A brick then touches
// two bricks of its own row and two of each neighbouring row, and colouring
// brick (i, j) with (i + 2 (j mod 2)) mod 3 separates it from all six: three
// colours, where a general colouring of breadth-first aggregates needs four or
// more.  Fails (the caller falls back) when a brick would exceed one
// wavefront of rows even after shrinking the bricks, or when a matrix entry
// couples two bricks of one colour (cells not small against a brick).
static bool brickAggregates(int n, const Graph &g, const double *xy, int target, const std::vector<int> &rpos,
                            std::vector<int> &blk, std::vector<std::vector<int>> &members,
                            std::vector<int> &colour) {
  double x0 = xy[0], x1 = xy[0], y0 = xy[1], y1 = xy[1];
  for (int i = 0; i < n; ++i) {
    x0 = std::min(x0, xy[2 * i]);
    x1 = std::max(x1, xy[2 * i]);
    y0 = std::min(y0, xy[2 * i + 1]);
    y1 = std::max(y1, xy[2 * i + 1]);
  }
  const double area = (x1 - x0) * (y1 - y0);
  if (!(area > 0.) || n < 4 * target) return false;
  std::vector<long long> key(n);
  std::vector<int> bi(n), bj(n), bk(n);  // bk: brick of each row (blk is written only on success)
  double side = std::sqrt(target * area / n);
  for (int attempt = 0;; ++attempt) {
    if (attempt == 6) return false;
    long long ni = 0;
    for (int r = 0; r < n; ++r) {
      const int j = (int)std::floor((xy[2 * r + 1] - y0) / side);
      const int i = (int)std::floor((xy[2 * r] - x0) / side - 0.5 * (j & 1)) + 1;
      bi[r] = i;
      bj[r] = j;
      ni = std::max<long long>(ni, i + 1);
    }
    for (int r = 0; r < n; ++r) key[r] = (long long)bj[r] * (ni + 1) + bi[r];
    std::vector<long long> uk(key);
    std::sort(uk.begin(), uk.end());
    uk.erase(std::unique(uk.begin(), uk.end()), uk.end());
    const int nb = (int)uk.size();
    std::vector<int> size(nb, 0);
    for (int r = 0; r < n; ++r) {
      bk[r] = (int)(std::lower_bound(uk.begin(), uk.end(), key[r]) - uk.begin());
      ++size[bk[r]];
    }
    if (*std::max_element(size.begin(), size.end()) > kLanes) {
      side *= 0.9;
      continue;
    }
    std::vector<int> bcol(nb);
    for (int r = 0; r < n; ++r) bcol[bk[r]] = (bi[r] + 2 * (bj[r] & 1)) % 3;
    for (int u = 0; u < n; ++u)
      for (int k = g.ptr[u]; k < g.ptr[u + 1]; ++k) {
        const int w = g.adj[k];
        if (bk[w] != bk[u] && bcol[bk[w]] == bcol[bk[u]]) return false;
      }
    // number the bricks by their first row in RCM order, as breadth-first
    // aggregation would have created them
    std::vector<int> first(nb, INT_MAX);
    for (int r = 0; r < n; ++r) first[bk[r]] = std::min(first[bk[r]], rpos[r]);
    std::vector<int> order(nb), renum(nb);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int c) { return first[a] < first[c]; });
    for (int k = 0; k < nb; ++k) renum[order[k]] = k;
    members.assign(nb, {});
    colour.assign(nb, 0);
    for (int k = 0; k < nb; ++k) colour[k] = bcol[order[k]];
    for (int r = 0; r < n; ++r) {
      blk[r] = renum[bk[r]];
      members[blk[r]].push_back(r);
    }
    for (auto &mem : members)
      std::sort(mem.begin(), mem.end(), [&](int a, int c) { return rpos[a] < rpos[c]; });
    return true;
  }
}

void buildOrdering(Ordering &o, int n, const int *rowPtr, const int *colInd, int blockRows, const double *cellXY,
                   bool blockJacobi, bool iluLimits) {
  if (n <= 0) fail(PHASE_HIP_EINVAL, "rank must be positive");
  if (blockRows < 1 || blockRows > kLanes) fail(PHASE_HIP_EINVAL, "blockRows must be in [1,64]");
  o = Ordering();
  o.n = n;
  // 1. compact pattern
  o.cRowPtr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    if (rowPtr[i + 1] < rowPtr[i]) fail(PHASE_HIP_EINVAL, "rowPtr is not monotone");
    for (int j = rowPtr[i]; j < rowPtr[i + 1]; ++j) {
      int c = colInd[j];
      if (c < 0) continue;
      if (c >= n) fail(PHASE_HIP_EINVAL, "column index out of range in row " + std::to_string(i));
      o.cCol.push_back(c);
      o.compactSrc.push_back(j);
    }
    o.cRowPtr[i + 1] = (int)o.cCol.size();
  }
  o.nnz = (int)o.cCol.size();

  // 2. RCM
  Graph g = symmetrise(n, o.cRowPtr, o.cCol);
  std::vector<int> rorder = rcm(g, n);
  std::vector<int> rpos(n);
  for (int k = 0; k < n; ++k) rpos[rorder[k]] = k;

  std::vector<int> blk(n, -1);
  std::vector<std::vector<int>> members;
  std::vector<int> colour;
  int nColours = 0;
  const bool bricks = cellXY && brickAggregates(n, g, cellXY, blockRows, rpos, blk, members, colour);
  if (!bricks) {
    // rows coupled to no other row (the solid cells of an immersed boundary)
    // would each be an aggregate of their own, and stay so on every coarser
    // level: they are grouped blockRows at a time in RCM order instead (they
    // couple to nothing, so a group of them factors exactly and coarsens to
    // another such row)
    std::vector<int> iso;
    for (int s0 : rorder)
      if (g.ptr[s0 + 1] == g.ptr[s0]) iso.push_back(s0);
    for (size_t k = 0; k < iso.size(); k += blockRows) {
      const int b = (int)members.size();
      members.emplace_back(iso.begin() + k, iso.begin() + std::min(iso.size(), k + blockRows));
      for (int r : members.back()) blk[r] = b;
    }
  }
  if (bricks) {
    nColours = 3;
  } else if (blockJacobi) {
    // breadth-first aggregates as below; no colouring (one colour)
    std::vector<int> queue, nbs;
    for (int s0 : rorder) {
      if (blk[s0] >= 0) continue;
      const int b = (int)members.size();
      members.emplace_back();
      auto &mem = members.back();
      mem.push_back(s0);
      blk[s0] = b;
      queue.assign(1, s0);
      size_t head = 0;
      while (head < queue.size() && (int)mem.size() < blockRows) {
        const int u = queue[head++];
        nbs.clear();
        for (int k = g.ptr[u]; k < g.ptr[u + 1]; ++k)
          if (blk[g.adj[k]] < 0) nbs.push_back(g.adj[k]);
        std::sort(nbs.begin(), nbs.end(), [&](int a, int c) { return rpos[a] < rpos[c]; });
        for (int w : nbs) {
          if ((int)mem.size() >= blockRows) break;
          blk[w] = b;
          mem.push_back(w);
          queue.push_back(w);
        }
      }
      std::sort(mem.begin(), mem.end(), [&](int a, int c) { return rpos[a] < rpos[c]; });
    }
    colour.assign(members.size(), 0);
    nColours = 1;
  } else {
    // 3. breadth-first aggregation into blocks of <= blockRows rows
    std::vector<int> queue, nbs;
    for (int s : rorder) {
      if (blk[s] >= 0) continue;
      int b = (int)members.size();
      members.emplace_back();
      auto &mem = members.back();
      mem.push_back(s);
      blk[s] = b;
      queue.assign(1, s);
      size_t head = 0;
      while (head < queue.size() && (int)mem.size() < blockRows) {
        int u = queue[head++];
        nbs.clear();
        for (int k = g.ptr[u]; k < g.ptr[u + 1]; ++k)
          if (blk[g.adj[k]] < 0) nbs.push_back(g.adj[k]);
        std::sort(nbs.begin(), nbs.end(), [&](int a, int c) { return rpos[a] < rpos[c]; });
        for (int w : nbs) {
          if ((int)mem.size() >= blockRows) break;
          blk[w] = b;
          mem.push_back(w);
          queue.push_back(w);
        }
      }
      std::sort(mem.begin(), mem.end(), [&](int a, int c) { return rpos[a] < rpos[c]; });
    }
    int nb = (int)members.size();

    // 4. colouring of the block graph, DSATUR (Brelaz 1979): repeatedly colour
    //    the block whose neighbours already use the most distinct colours (ties:
    //    most neighbours, then lowest index) with the smallest free colour.
    //    Every colour costs one kernel launch per sweep, so fewer is faster.
    std::vector<std::vector<int>> badj(nb);
    for (int b = 0; b < nb; ++b) {
      auto &a = badj[b];
      for (int u : members[b])
        for (int k = g.ptr[u]; k < g.ptr[u + 1]; ++k) {
          int ob = blk[g.adj[k]];
          if (ob != b) a.push_back(ob);
        }
      std::sort(a.begin(), a.end());
      a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    colour.assign(nb, -1);
    std::vector<int> satur(nb, 0);
    std::vector<uint64_t> usedMask(nb, 0);  // colours seen among neighbours (< 64)
    // priority: (saturation, degree, -index); max-heap with lazy entries
    using Key = std::tuple<int, int, int>;
    std::priority_queue<Key> heap;
    for (int b = 0; b < nb; ++b) heap.push(Key(0, (int)badj[b].size(), -b));
    while (!heap.empty()) {
      auto [sat, deg, nbIdx] = heap.top();
      heap.pop();
      const int b = -nbIdx;
      if (colour[b] >= 0 || sat != satur[b]) continue;  // stale entry
      int c = 0;
      while (c < 64 && (usedMask[b] >> c & 1ull)) ++c;
      if (c >= 64) fail(PHASE_HIP_EINVAL, "block graph needs more than 64 colours");
      colour[b] = c;
      nColours = std::max(nColours, c + 1);
      for (int ob : badj[b]) {
        if (colour[ob] >= 0 || (usedMask[ob] >> c & 1ull)) continue;
        usedMask[ob] |= 1ull << c;
        ++satur[ob];
        heap.push(Key(satur[ob], (int)badj[ob].size(), -ob));
      }
    }
    // 4b. fewer colours where Kempe-chain interchanges, finished by a bounded
    //     Tabucol search for the blocks they leave, find them; each colour
    //     removed saves two launches per preconditioner apply.
    while (nColours > 2) {
      std::vector<int> trial = colour;
      const int maxLeft = std::max(16, nb / 1000);
      const int left = kempeRecolour(badj, nColours, trial, maxLeft);
      if (left > 0 && (left > maxLeft || !tabuRecolour(badj, nColours - 1, trial, 100000))) break;
      colour.swap(trial);
      --nColours;
    }
  }
  if (blockJacobi) {
    colour.assign(members.size(), 0);
    nColours = 1;
  }
  int nb = (int)members.size();
  o.nColours = nColours;

  // 5. ILU blocks: aggregates of one colour share no coupling, so consecutive
  //    aggregates of a colour are packed into one block of <= 64 rows (one
  //    wavefront).  The block's in-block dependency graph then has one
  //    component per aggregate and its level count is the largest of theirs:
  //    smaller aggregates cut the levels a wavefront walks, at the cost of
  //    more coupling between blocks.  Solver order = blocks by colour.
  std::vector<int> border(nb);
  std::iota(border.begin(), border.end(), 0);
  std::stable_sort(border.begin(), border.end(), [&](int a, int c) { return colour[a] < colour[c]; });
  std::vector<std::vector<int>> packs;
  std::vector<int> packColour;
  std::vector<int> aggSize;  // aggregates in pack order, i.e. in solver order
  for (int ci = 0; ci < nb; ++ci) {
    const int a = border[ci];
    // block Jacobi keeps one aggregate per block: couplings between the rows of
    // packed aggregates would lengthen the block's dependency levels
    if (packs.empty() || packColour.back() != colour[a] || blockJacobi ||
        packs.back().size() + members[a].size() > (size_t)kLanes) {
      packs.emplace_back();
      packColour.push_back(colour[a]);
    }
    packs.back().insert(packs.back().end(), members[a].begin(), members[a].end());
    aggSize.push_back((int)members[a].size());
  }
  o.aggPtr.assign(aggSize.size() + 1, 0);
  for (size_t a = 0; a < aggSize.size(); ++a) o.aggPtr[a + 1] = o.aggPtr[a] + aggSize[a];
  nb = (int)packs.size();
  o.perm.clear();
  o.perm.reserve(n);
  o.nBlocks = nb;
  o.blocks.assign(nb, BlockInfo{});
  std::vector<int> blockOfRow(n);
  for (int bi = 0; bi < nb; ++bi) {
    auto &mem = packs[bi];
    o.blocks[bi].rowStart = (int)o.perm.size();
    o.blocks[bi].nRows = (int)mem.size();
    for (int r : mem) {
      blockOfRow[o.perm.size()] = bi;
      o.perm.push_back(r);
    }
  }
  o.iperm.assign(n, 0);
  for (int k = 0; k < n; ++k) o.iperm[o.perm[k]] = k;
  o.colourPtr.assign(nColours + 1, 0);
  for (int b = 0; b < nb; ++b) ++o.colourPtr[packColour[b] + 1];
  for (int c = 0; c < nColours; ++c) o.colourPtr[c + 1] += o.colourPtr[c];

  // 6. SELL-64 layout of A over solver rows, original in-row order
  o.nSlices = (n + kLanes - 1) / kLanes;
  o.sliceOff.assign(o.nSlices + 1, 0);
  for (int s = 0; s < o.nSlices; ++s) {
    int K = 0;
    for (int l = 0; l < kLanes && s * kLanes + l < n; ++l) {
      int i = o.perm[s * kLanes + l];
      K = std::max(K, o.cRowPtr[i + 1] - o.cRowPtr[i]);
    }
    o.sliceOff[s + 1] = o.sliceOff[s] + K * kLanes;
  }
  o.sellCol.assign(o.sliceOff[o.nSlices], -1);
  o.sellSrc.assign(o.sliceOff[o.nSlices], -1);
  for (int k = 0; k < n; ++k) {
    int s = k / kLanes, l = k % kLanes, i = o.perm[k];
    for (int e = 0, j = o.cRowPtr[i]; j < o.cRowPtr[i + 1]; ++j, ++e) {
      int idx = o.sliceOff[s] + e * kLanes + l;
      o.sellCol[idx] = o.iperm[o.cCol[j]];
      o.sellSrc[idx] = o.compactSrc[j];
    }
  }

  // 7. ILU(0) factor layout
  struct Ent { int col, src; };
  std::vector<std::vector<Ent>> srow(n);
  o.rows.assign(n, RowInfo{});
  std::vector<int> diagSrc(n, -1);
  for (int k = 0; k < n; ++k) {
    int i = o.perm[k];
    auto &r = srow[k];
    const BlockInfo &Bk = o.blocks[blockOfRow[k]];
    for (int j = o.cRowPtr[i]; j < o.cRowPtr[i + 1]; ++j) {
      const int c = o.iperm[o.cCol[j]];
      // block Jacobi: the factor keeps the couplings inside the row's block only
      if (blockJacobi && (c < Bk.rowStart || c >= Bk.rowStart + Bk.nRows)) continue;
      r.push_back({c, o.compactSrc[j]});
    }
    std::sort(r.begin(), r.end(), [](const Ent &a, const Ent &b) { return a.col < b.col; });
    // merge duplicate columns (possible only in hand-made patterns): keep first
    std::vector<Ent> u;
    for (auto &e : r)
      if (u.empty() || u.back().col != e.col) u.push_back(e);
    r.swap(u);
  }
  int lRun = 0, uRun = 0, maxRegs = 0;
  for (int b = 0; b < nb; ++b) {
    BlockInfo &B = o.blocks[b];
    int bs = B.rowStart, be = bs + B.nRows;
    int kL = 0, kU = 0, nLev = 0, nULev = 0;
    for (int k = bs; k < be; ++k) {
      int nL = 0, nLoff = 0, nU = 0, nUin = 0, lev = 0;
      for (auto &e : srow[k]) {
        if (e.col < k) {
          ++nL;
          if (e.col < bs) ++nLoff;
          else lev = std::max(lev, (int)o.rows[e.col].lev + 1);
        } else if (e.col > k) {
          ++nU;
          if (e.col < be) ++nUin;
        } else {
          diagSrc[k] = e.src;
        }
      }
      if (nL > 255 || nU > 255) fail(PHASE_HIP_EINVAL, "row too long for the ILU layout");
      // the lane record holds these counts in 4 bits (the sweep kernels keep
      // up to 8 in registers and read the rest from the factor); an ordering
      // whose ILU(0) is never applied (a Jacobi-smoothed multigrid level)
      // takes any width, its lane records unused
      if (iluLimits && nU - nUin > 15)
        fail(PHASE_HIP_EINVAL, "row couples to more than 15 rows of later ILU blocks");
      if (iluLimits && (nL - nLoff > 15 || nUin > 15))
        fail(PHASE_HIP_EINVAL, "row couples to more than 15 rows of its own ILU block");
      if (iluLimits && nLoff > 63) fail(PHASE_HIP_EINVAL, "row couples to more than 63 rows of earlier ILU blocks");
      maxRegs = std::max({maxRegs, nL - nLoff, nLoff, nUin, nU - nUin});
      o.rows[k].nL = (uint8_t)nL;
      o.rows[k].nLoff = (uint8_t)nLoff;
      o.rows[k].nU = (uint8_t)nU;
      o.rows[k].nUin = (uint8_t)nUin;
      o.rows[k].lev = (uint8_t)lev;
      kL = std::max(kL, nL);
      kU = std::max(kU, nU);
      nLev = std::max(nLev, lev + 1);
    }
    for (int k = be - 1; k >= bs; --k) {
      int ulev = 0;
      for (auto &e : srow[k])
        if (e.col > k && e.col < be) ulev = std::max(ulev, (int)o.rows[e.col].ulev + 1);
      o.rows[k].ulev = (uint8_t)ulev;
      nULev = std::max(nULev, ulev + 1);
    }
    B.lOff = lRun;
    B.kL = kL;
    B.uOff = uRun;
    B.kU = kU;
    B.nLev = nLev;
    B.nULev = nULev;
    lRun += kL * kLanes;
    uRun += kU * kLanes;
  }
  o.nLSlots = lRun;
  o.regWidth = maxRegs <= 4 ? 4 : 8;
  o.nUSlots = uRun;
  o.lanes.assign((size_t)nb * kLanes, LaneRec{-1, 0u});
  for (int b = 0; b < nb; ++b) {
    const BlockInfo &B = o.blocks[b];
    for (int l = 0; l < B.nRows; ++l) {
      const int k = B.rowStart + l;
      const RowInfo &r = o.rows[k];
      o.lanes[(size_t)b * kLanes + l] =
          LaneRec{k, packLane(std::min(r.nL - r.nLoff, 15), std::min((int)r.nLoff, 63), std::min((int)r.nUin, 15),
                              std::min(r.nU - r.nUin, 15), r.lev, r.ulev)};
    }
  }
  o.lCol.assign(lRun, -1);
  o.uCol.assign(uRun, -1);
  o.facSrc.assign((size_t)lRun + uRun + n, -1);
  // fac index of entry (row k, column c), or -1
  auto facIndex = [&](int k, int c) -> int {
    const BlockInfo &B = o.blocks[blockOfRow[k]];
    int lane = k - B.rowStart, eL = 0, eU = 0;
    for (auto &e : srow[k]) {
      if (e.col < k) {
        if (e.col == c) return B.lOff + eL * kLanes + lane;
        ++eL;
      } else if (e.col > k) {
        if (e.col == c) return lRun + B.uOff + eU * kLanes + lane;
        ++eU;
      } else if (c == k) {
        return lRun + uRun + k;
      }
    }
    return -1;
  };
  for (int k = 0; k < n; ++k) {
    const BlockInfo &B = o.blocks[blockOfRow[k]];
    int lane = k - B.rowStart, eL = 0, eU = 0;
    for (auto &e : srow[k]) {
      if (e.col < k) {
        int s = B.lOff + eL * kLanes + lane;
        o.lCol[s] = e.col;
        o.facSrc[s] = e.src;
        ++eL;
      } else if (e.col > k) {
        int s = B.uOff + eU * kLanes + lane;
        o.uCol[s] = e.col;
        o.facSrc[lRun + s] = e.src;
        ++eU;
      }
    }
    o.facSrc[(size_t)lRun + uRun + k] = diagSrc[k];
  }
  // pred / succ block lists
  std::vector<int> tmp;
  for (int b = 0; b < nb; ++b) {
    BlockInfo &B = o.blocks[b];
    int bs = B.rowStart, be = bs + B.nRows;
    tmp.clear();
    for (int k = bs; k < be; ++k)
      for (auto &e : srow[k])
        if (e.col < bs) tmp.push_back(blockOfRow[e.col]);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    B.predBeg = (int)o.pred.size();
    o.pred.insert(o.pred.end(), tmp.begin(), tmp.end());
    B.predEnd = (int)o.pred.size();
    tmp.clear();
    for (int k = bs; k < be; ++k)
      for (auto &e : srow[k])
        if (e.col >= be) tmp.push_back(blockOfRow[e.col]);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    B.succBeg = (int)o.succ.size();
    o.succ.insert(o.succ.end(), tmp.begin(), tmp.end());
    B.succEnd = (int)o.succ.size();
    if (B.predEnd - B.predBeg > kLanes || B.succEnd - B.succBeg > kLanes)
      fail(PHASE_HIP_EINVAL, "block has more than 64 neighbour blocks");
  }
  // IKJ update lists: for L entry (k, c) in ascending c, every U entry (c, j)
  // of row c whose column j is also in row k, in ascending j.
  auto visitPairs = [&](auto &&emit) {
    for (int k = 0; k < n; ++k) {
      const BlockInfo &B = o.blocks[blockOfRow[k]];
      int lane = k - B.rowStart, eL = 0;
      for (auto &e : srow[k]) {
        if (e.col >= k) break;
        int slot = B.lOff + eL * kLanes + lane;
        ++eL;
        int c = e.col;
        const BlockInfo &C = o.blocks[blockOfRow[c]];
        int laneC = c - C.rowStart, eU = 0;
        for (auto &f : srow[c]) {
          if (f.col <= c) continue;
          int src = lRun + C.uOff + eU * kLanes + laneC;
          ++eU;
          int dst = facIndex(k, f.col);
          if (dst >= 0) emit(slot, src, dst);
        }
      }
    }
  };
  o.updPtr.assign(lRun + 1, 0);
  visitPairs([&](int slot, int, int) { ++o.updPtr[slot + 1]; });
  for (int s = 0; s < lRun; ++s) o.updPtr[s + 1] += o.updPtr[s];
  o.updSrc.resize(o.updPtr[lRun]);
  o.updDst.resize(o.updPtr[lRun]);
  std::vector<int> fillPos(o.updPtr.begin(), o.updPtr.end() - 1);
  visitPairs([&](int slot, int src, int dst) {
    o.updSrc[fillPos[slot]] = src;
    o.updDst[fillPos[slot]++] = dst;
  });
}

}  // namespace phase

namespace phase {

void buildCoarse(const Ordering &fine, const double *fineXY, CoarseLevel &c) {
  const int n = fine.n;
  const int na = (int)fine.aggPtr.size() - 1;
  if (na < 1 || fine.aggPtr[na] != n) fail(PHASE_HIP_EINVAL, "ordering has no aggregates");
  c = CoarseLevel();
  c.n = na;
  c.aggOfRow.assign(n, 0);
  for (int a = 0; a < na; ++a)
    for (int k = fine.aggPtr[a]; k < fine.aggPtr[a + 1]; ++k) c.aggOfRow[k] = a;
  // coarse entry (A, B) of every finer compact entry, with its caller slot
  struct Ent {
    int col, slot;
  };
  std::vector<std::vector<Ent>> rows(na);
  for (int i = 0; i < n; ++i) {
    const int A = c.aggOfRow[fine.iperm[i]];
    for (int j = fine.cRowPtr[i]; j < fine.cRowPtr[i + 1]; ++j)
      rows[A].push_back({c.aggOfRow[fine.iperm[fine.cCol[j]]], fine.compactSrc[j]});
  }
  c.rowPtr.assign(na + 1, 0);
  c.galPtr.clear();
  for (int A = 0; A < na; ++A) {
    auto &r = rows[A];
    std::sort(r.begin(), r.end(),
              [](const Ent &x, const Ent &y) { return x.col < y.col || (x.col == y.col && x.slot < y.slot); });
    for (size_t q = 0; q < r.size(); ++q) {
      if (q == 0 || r[q].col != r[q - 1].col) {
        c.colInd.push_back(r[q].col);
        c.galPtr.push_back((int)c.galSrc.size());
      }
      c.galSrc.push_back(r[q].slot);
    }
    c.rowPtr[A + 1] = (int)c.colInd.size();
  }
  c.galPtr.push_back((int)c.galSrc.size());
  if (fineXY) {
    c.xy.assign(2 * (size_t)na, 0.);
    for (int a = 0; a < na; ++a) {
      double sx = 0., sy = 0.;
      for (int k = fine.aggPtr[a]; k < fine.aggPtr[a + 1]; ++k) {
        sx += fineXY[2 * fine.perm[k]];
        sy += fineXY[2 * fine.perm[k] + 1];
      }
      const double m = fine.aggPtr[a + 1] - fine.aggPtr[a];
      c.xy[2 * a] = sx / m;
      c.xy[2 * a + 1] = sy / m;
    }
  }
}

}  // namespace phase
