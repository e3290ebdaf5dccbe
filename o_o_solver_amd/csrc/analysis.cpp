// Host analysis ahead of the numeric factorisation: a fill-reducing node ordering (geometric nested dissection) and
// the symbolic factor (elimination tree, postorder, column counts, fundamental and relaxed supernodes, row
// structures, sizes and flops). These stay on the host; the numeric phase on the GPU consumes their result.
//
// The symbolic analysis computes the structure the reference defines in SymbolicFactor::analyse
// (src/linalg/symbolic/SymbolicFactor.cpp:166-296): the elimination tree of the given order
// (src/linalg/symbolic/EliminationTree.cpp:12-36), the order postordered with children and roots in increasing
// position (SymbolicFactor.cpp:49-99), the fundamental supernodes (src/linalg/symbolic/Supernodes.cpp:30-70), the
// relaxed amalgamation with the thresholds {24, 96, 288} / {0.8, 0.1, 0.05} (src/linalg/symbolic/Amalgamation.cpp:11-88,
// include/oos/linalg/symbolic/Amalgamation.hpp:22-32), the row structure of each supernode and the flop count
// (SymbolicFactor.cpp:222-295). For the same graph, weights and order the structure is the same, bit for bit: the
// tests compare it against the compiled reference. The algorithms are this project's own (column counts by row
// subtrees, for instance, where the reference uses Gilbert-Ng-Peyton).
//
// The ordering is this project's own nested dissection on node coordinates: a region is cut at the weighted median
// of its largest extent, the separator being a minimum vertex cover of the cut edges (König: from a maximum
// matching), numbered after both halves; disconnected regions are split into components; leaves of at most
// LEAF_NODES nodes are ordered by exact minimum degree. It is not the reference's GeometricDissection (which moves
// the cut to a change of coordinate, weights the cover by a maximum flow and orders the leaves by constrained AMD),
// so its permutation differs; the bench reports the flops of both orderings.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{

constexpr std::int32_t ROOT = -1;

// ---- nested dissection -------------------------------------------------------------------------------------------

struct Dissection
{
    std::int32_t n;
    const std::int64_t* off;
    const std::int32_t* adj;
    const std::int32_t* weight;
    const double* xyz;
    std::int32_t leafNodes;
    std::vector<std::int32_t> order;
    std::vector<std::int32_t> region; // region id of each node (stamp), -1 outside
    std::vector<std::int32_t> side;   // scratch: left / right / separator marks of the current cut
    std::vector<std::int32_t> seen;   // scratch: the component search's visit stamps
    std::int32_t stamp = 0;

    void orderLeaf(std::vector<std::int32_t>& nodes)
    {
        // Exact minimum degree on the induced graph (explicit elimination graph; leaves are small).
        const std::size_t m = nodes.size();
        if (m == 0)
            return;
        std::sort(nodes.begin(), nodes.end());
        std::vector<std::vector<char>> link(m, std::vector<char>(m, 0));
        for (std::size_t a = 0; a < m; ++a)
            for (std::int64_t k = off[nodes[a]]; k < off[nodes[a] + 1]; ++k)
            {
                const auto it = std::lower_bound(nodes.begin(), nodes.end(), adj[k]);
                if (it != nodes.end() && *it == adj[k])
                    link[a][static_cast<std::size_t>(it - nodes.begin())] = 1;
            }
        std::vector<char> done(m, 0);
        for (std::size_t step = 0; step < m; ++step)
        {
            std::size_t best = m;
            int bestDegree = 1 << 30;
            for (std::size_t a = 0; a < m; ++a)
            {
                if (done[a])
                    continue;
                int degree = 0;
                for (std::size_t b = 0; b < m; ++b)
                    degree += (!done[b] && link[a][b]) ? 1 : 0;
                if (degree < bestDegree)
                {
                    bestDegree = degree;
                    best = a;
                }
            }
            done[best] = 1;
            order.push_back(nodes[best]);
            // Its neighbours become a clique.
            std::vector<std::size_t> nb;
            for (std::size_t b = 0; b < m; ++b)
                if (!done[b] && link[best][b])
                    nb.push_back(b);
            for (const std::size_t a : nb)
                for (const std::size_t b : nb)
                    if (a != b)
                        link[a][b] = 1;
        }
    }

    // Connected components of a region (nodes stamped with `id`), each sorted.
    std::vector<std::vector<std::int32_t>> components(const std::vector<std::int32_t>& nodes, std::int32_t id)
    {
        std::vector<std::vector<std::int32_t>> result;
        std::vector<std::int32_t> stack;
        for (const std::int32_t start : nodes)
        {
            if (seen[start] == id)
                continue;
            std::vector<std::int32_t> comp;
            seen[start] = id;
            stack.push_back(start);
            while (!stack.empty())
            {
                const std::int32_t v = stack.back();
                stack.pop_back();
                comp.push_back(v);
                for (std::int64_t k = off[v]; k < off[v + 1]; ++k)
                {
                    const std::int32_t u = adj[k];
                    if (region[u] == id && seen[u] != id)
                    {
                        seen[u] = id;
                        stack.push_back(u);
                    }
                }
            }
            std::sort(comp.begin(), comp.end());
            result.push_back(std::move(comp));
        }
        return result;
    }

    void dissect(std::vector<std::int32_t> nodes)
    {
        if (static_cast<std::int32_t>(nodes.size()) <= leafNodes)
        {
            orderLeaf(nodes);
            return;
        }
        const std::int32_t id = ++stamp;
        for (const std::int32_t v : nodes)
            region[v] = id;
        auto comps = components(nodes, id);
        if (comps.size() > 1)
        {
            for (auto& comp : comps)
                dissect(std::move(comp));
            return;
        }
        // Cut at the weighted median of the largest extent.
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (const std::int32_t v : nodes)
            for (int a = 0; a < 3; ++a)
            {
                lo[a] = std::min(lo[a], xyz[3 * v + a]);
                hi[a] = std::max(hi[a], xyz[3 * v + a]);
            }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (hi[a] - lo[a] > hi[axis] - lo[axis])
                axis = a;
        std::sort(nodes.begin(), nodes.end(),
                  [&](std::int32_t a, std::int32_t b)
                  {
                      const double ca = xyz[3 * a + axis], cb = xyz[3 * b + axis];
                      return ca < cb || (ca == cb && a < b);
                  });
        std::int64_t total = 0;
        for (const std::int32_t v : nodes)
            total += weight[v];
        std::int64_t acc = 0;
        std::size_t split = 0;
        while (split < nodes.size() && 2 * (acc + weight[nodes[split]]) <= total)
            acc += weight[nodes[split++]];
        split = std::clamp<std::size_t>(split, 1, nodes.size() - 1);
        // Move the cut to the nearest change of coordinate, so that a line of a structured mesh is not split.
        auto coordAt = [&](std::size_t k) { return xyz[3 * nodes[k] + axis]; };
        if (coordAt(split - 1) == coordAt(split))
        {
            std::size_t down = split, up = split;
            while (down > 1 && coordAt(down - 1) == coordAt(down))
                --down;
            while (up + 1 < nodes.size() && coordAt(up - 1) == coordAt(up))
                ++up;
            const bool downOk = coordAt(down - 1) != coordAt(down), upOk = coordAt(up - 1) != coordAt(up);
            if (downOk && (!upOk || split - down <= up - split))
                split = down;
            else if (upOk)
                split = up;
        }
        // side: id·2 left, id·2+1 right (stamps are unique per call).
        const std::int32_t leftMark = 2 * id, rightMark = 2 * id + 1;
        for (std::size_t k = 0; k < nodes.size(); ++k)
            side[nodes[k]] = k < split ? leftMark : rightMark;
        // Bipartite graph of the cut edges: left nodes with a right neighbour, right nodes with a left neighbour.
        std::vector<std::int32_t> U, V;
        for (const std::int32_t v : nodes)
        {
            bool cut = false;
            for (std::int64_t k = off[v]; k < off[v + 1] && !cut; ++k)
            {
                const std::int32_t u = adj[k];
                cut = region[u] == id && side[u] != side[v];
            }
            if (cut)
                (side[v] == leftMark ? U : V).push_back(v);
        }
        std::sort(U.begin(), U.end());
        std::sort(V.begin(), V.end());
        auto indexIn = [](const std::vector<std::int32_t>& list, std::int32_t v)
        {
            const auto it = std::lower_bound(list.begin(), list.end(), v);
            return (it != list.end() && *it == v) ? static_cast<std::int32_t>(it - list.begin()) : -1;
        };
        std::vector<std::vector<std::int32_t>> edges(U.size());
        for (std::size_t a = 0; a < U.size(); ++a)
            for (std::int64_t k = off[U[a]]; k < off[U[a] + 1]; ++k)
            {
                const std::int32_t b = indexIn(V, adj[k]);
                if (b >= 0)
                    edges[a].push_back(b);
            }
        // Maximum matching (Kuhn's augmenting paths, iterative).
        std::vector<std::int32_t> matchU(U.size(), -1), matchV(V.size(), -1), visit(V.size(), -1);
        for (std::size_t a0 = 0; a0 < U.size(); ++a0)
        {
            // DFS with an explicit stack of (u, next edge).
            std::vector<std::pair<std::int32_t, std::size_t>> stack{{static_cast<std::int32_t>(a0), 0}};
            std::vector<std::int32_t> via; // the v taken at each level
            bool found = false;
            while (!stack.empty() && !found)
            {
                auto& [a, e] = stack.back();
                if (e >= edges[a].size())
                {
                    stack.pop_back();
                    if (!via.empty())
                        via.pop_back();
                    continue;
                }
                const std::int32_t b = edges[a][e++];
                if (visit[b] == static_cast<std::int32_t>(a0))
                    continue;
                visit[b] = static_cast<std::int32_t>(a0);
                via.push_back(b);
                if (matchV[b] < 0)
                {
                    found = true;
                    break;
                }
                stack.push_back({matchV[b], 0});
            }
            if (found)
            {
                // Augment along the stack: stack[i].first matched to via[i].
                for (std::size_t i = 0; i < via.size(); ++i)
                {
                    const std::int32_t a = stack[i].first, b = via[i];
                    matchU[a] = b;
                    matchV[b] = a;
                }
            }
        }
        // König: Z = reachable from unmatched U by alternating paths; cover = (U \ Z) ∪ (V ∩ Z).
        std::vector<char> zU(U.size(), 0), zV(V.size(), 0);
        std::vector<std::int32_t> queue;
        for (std::size_t a = 0; a < U.size(); ++a)
            if (matchU[a] < 0)
            {
                zU[a] = 1;
                queue.push_back(static_cast<std::int32_t>(a));
            }
        for (std::size_t q = 0; q < queue.size(); ++q)
        {
            const std::int32_t a = queue[q];
            for (const std::int32_t b : edges[a])
                if (!zV[b] && matchU[a] != b)
                {
                    zV[b] = 1;
                    const std::int32_t next = matchV[b];
                    if (next >= 0 && !zU[next])
                    {
                        zU[next] = 1;
                        queue.push_back(next);
                    }
                }
        }
        std::vector<std::int32_t> separator;
        for (std::size_t a = 0; a < U.size(); ++a)
            if (!zU[a])
                separator.push_back(U[a]);
        for (std::size_t b = 0; b < V.size(); ++b)
            if (zV[b])
                separator.push_back(V[b]);
        const std::int32_t sepMark = -2 - id;
        for (const std::int32_t v : separator)
            side[v] = sepMark;
        std::vector<std::int32_t> left, right;
        for (const std::int32_t v : nodes)
        {
            if (side[v] == leftMark)
                left.push_back(v);
            else if (side[v] == rightMark)
                right.push_back(v);
        }
        std::sort(left.begin(), left.end());
        std::sort(right.begin(), right.end());
        // The separator along the cut line: by the coordinate of the second largest extent.
        int along = axis == 0 ? 1 : 0;
        for (int a = 0; a < 3; ++a)
            if (a != axis && hi[a] - lo[a] > hi[along] - lo[along])
                along = a;
        std::sort(separator.begin(), separator.end(),
                  [&](std::int32_t a, std::int32_t b)
                  {
                      const double ca = xyz[3 * a + along], cb = xyz[3 * b + along];
                      return ca < cb || (ca == cb && a < b);
                  });
        for (const std::int32_t v : nodes)
            region[v] = -1;
        dissect(std::move(left));
        dissect(std::move(right));
        for (const std::int32_t v : separator)
            order.push_back(v);
    }
};

// ---- symbolic analysis -------------------------------------------------------------------------------------------

struct Thresholds
{
    std::int64_t columnLimits[3] = {24, 96, 288};
    double zeroFractions[3] = {0.8, 0.1, 0.05};
};

struct Symbolic
{
    std::int32_t nodeCount = 0;
    std::int32_t dofCount = 0;
    std::vector<std::int32_t> nodeOrder;         // node at each final position
    std::vector<std::int32_t> firstColumnOfPos;  // nodeCount + 1
    std::vector<std::int32_t> superFirstPos;     // supernodes + 1
    std::vector<std::int32_t> superParent;       // supernodes
    std::vector<std::int64_t> rowPosStart;       // supernodes + 1
    std::vector<std::int32_t> rowPos;
    std::vector<std::int64_t> rowStart;          // supernodes + 1
    std::vector<std::int32_t> rows;
    std::int32_t trailingDofCount = 0;
    std::int64_t panelValueCount = 0, factorEntryCount = 0, exactEntryCount = 0, largestFrontValues = 0,
                 updateStackPeakValues = 0;
    double flopCount = 0.0;
};

// Liu's elimination tree of the graph under `order` (node at each position), parent by position.
std::vector<std::int32_t> eliminationTree(std::int32_t n, const std::int64_t* off, const std::int32_t* adj,
                                          const std::vector<std::int32_t>& order, const std::vector<std::int32_t>& pos)
{
    std::vector<std::int32_t> parent(static_cast<std::size_t>(n), ROOT), ancestor(static_cast<std::size_t>(n), ROOT);
    for (std::int32_t k = 0; k < n; ++k)
    {
        const std::int32_t node = order[k];
        for (std::int64_t e = off[node]; e < off[node + 1]; ++e)
        {
            std::int32_t i = pos[adj[e]];
            while (i != ROOT && i < k)
            {
                const std::int32_t next = ancestor[i];
                ancestor[i] = k;
                if (next == ROOT)
                    parent[i] = k;
                i = next;
            }
        }
    }
    return parent;
}

// The positions [0, interior) in a depth-first postorder (children and roots in increasing position); a position whose
// parent is ROOT or a trailing position is a root. Returns the new node order.
std::vector<std::int32_t> postorderInterior(const std::vector<std::int32_t>& order,
                                            const std::vector<std::int32_t>& parent, std::int32_t interior)
{
    const std::size_t n = order.size();
    std::vector<std::vector<std::int32_t>> kids(static_cast<std::size_t>(interior));
    auto isRoot = [&](std::int32_t j) { return parent[j] == ROOT || parent[j] >= interior; };
    for (std::int32_t j = 0; j < interior; ++j)
        if (!isRoot(j))
            kids[parent[j]].push_back(j); // increasing j
    std::vector<std::int32_t> result;
    result.reserve(n);
    std::vector<std::pair<std::int32_t, std::size_t>> stack;
    for (std::int32_t root = 0; root < interior; ++root)
    {
        if (!isRoot(root))
            continue;
        stack.push_back({root, 0});
        while (!stack.empty())
        {
            auto& [p, next] = stack.back();
            if (next < kids[p].size())
            {
                const std::int32_t child = kids[p][next++];
                stack.push_back({child, 0});
            }
            else
            {
                result.push_back(order[p]);
                stack.pop_back();
            }
        }
    }
    for (std::size_t j = static_cast<std::size_t>(interior); j < n; ++j)
        result.push_back(order[j]);
    return result;
}

bool mergeIsWorthIt(const Thresholds& t, std::int64_t columns, std::int64_t zeros, std::int64_t rowsBelow)
{
    if (columns <= t.columnLimits[0])
        return true;
    const double entries = static_cast<double>(columns) * static_cast<double>(columns + 1) / 2.0 +
                           static_cast<double>(columns) * static_cast<double>(rowsBelow);
    const double fraction = static_cast<double>(zeros) / entries;
    return (columns <= t.columnLimits[1] && fraction < t.zeroFractions[0]) ||
           (columns <= t.columnLimits[2] && fraction < t.zeroFractions[1]) || fraction < t.zeroFractions[2];
}

Symbolic analyse(std::int32_t n, const std::int64_t* off, const std::int32_t* adj, const std::int32_t* weights,
                 const std::int32_t* givenOrder, std::int32_t trailing, bool amalgamate)
{
    Symbolic result;
    result.nodeCount = n;
    const std::int32_t interior = n - trailing;
    std::vector<std::int32_t> order(givenOrder, givenOrder + n), pos(static_cast<std::size_t>(n));
    for (std::int32_t p = 0; p < n; ++p)
        pos[order[p]] = p;
    {
        const auto given = eliminationTree(n, off, adj, order, pos);
        order = postorderInterior(order, given, interior);
    }
    for (std::int32_t p = 0; p < n; ++p)
        pos[order[p]] = p;
    const std::vector<std::int32_t> parent = eliminationTree(n, off, adj, order, pos);

    // Column counts by row subtrees: row k lies in column j < k exactly when j is on the tree path from a
    // neighbour i < k of k up to k.
    std::vector<std::int32_t> weightAt(static_cast<std::size_t>(n));
    for (std::int32_t p = 0; p < n; ++p)
        weightAt[p] = weights[order[p]];
    std::vector<std::int32_t> belowNodes(static_cast<std::size_t>(n), 0), mark(static_cast<std::size_t>(n), -1);
    std::vector<std::int64_t> belowWeight(static_cast<std::size_t>(n), 0);
    for (std::int32_t k = 0; k < n; ++k)
    {
        mark[k] = k;
        const std::int32_t node = order[k];
        for (std::int64_t e = off[node]; e < off[node + 1]; ++e)
        {
            std::int32_t i = pos[adj[e]];
            if (i >= k)
                continue;
            while (mark[i] != k)
            {
                mark[i] = k;
                ++belowNodes[i];
                belowWeight[i] += weightAt[k];
                i = parent[i];
            }
        }
    }

    result.firstColumnOfPos.assign(static_cast<std::size_t>(n) + 1, 0);
    for (std::int32_t k = 0; k < n; ++k)
    {
        result.firstColumnOfPos[k + 1] = result.firstColumnOfPos[k] + weightAt[k];
        const std::int64_t w = weightAt[k];
        result.exactEntryCount += w * (w + 1) / 2 + w * belowWeight[k];
        if (k >= interior)
            result.trailingDofCount += weightAt[k];
    }
    result.dofCount = result.firstColumnOfPos[n];

    // Fundamental supernodes.
    const std::int32_t trailingStart = trailing > 0 ? interior : n;
    std::vector<std::int32_t> childCount(static_cast<std::size_t>(n), 0);
    for (std::int32_t j = 0; j < n; ++j)
        if (parent[j] != ROOT)
            ++childCount[parent[j]];
    std::vector<std::int32_t> superOf(static_cast<std::size_t>(n));
    std::int32_t count = 0;
    for (std::int32_t j = 0; j < n; ++j)
    {
        bool continues = false;
        if (j > 0 && j >= trailingStart)
            continues = j > trailingStart;
        else if (j > 0)
            continues = parent[j - 1] == j && childCount[j] == 1 && belowNodes[j - 1] == belowNodes[j] + 1;
        if (!continues)
            ++count;
        superOf[j] = count - 1;
    }
    std::vector<std::int32_t> fFirst(static_cast<std::size_t>(count) + 1), fParent(static_cast<std::size_t>(count));
    for (std::int32_t j = n - 1; j >= 0; --j)
        fFirst[superOf[j]] = j;
    fFirst[count] = n;
    for (std::int32_t s = 0; s < count; ++s)
    {
        const std::int32_t up = parent[fFirst[s + 1] - 1];
        fParent[s] = up == ROOT ? ROOT : superOf[up];
    }

    // Relaxed amalgamation (children merged into the parent just after them, from the top down).
    Thresholds thresholds;
    if (!amalgamate)
    {
        thresholds.columnLimits[0] = thresholds.columnLimits[1] = thresholds.columnLimits[2] = 0;
        thresholds.zeroFractions[0] = thresholds.zeroFractions[1] = thresholds.zeroFractions[2] = -1.0;
    }
    std::vector<std::int64_t> cols(static_cast<std::size_t>(count), 0), below(static_cast<std::size_t>(count)),
        zeros(static_cast<std::size_t>(count), 0);
    std::vector<std::int32_t> top(static_cast<std::size_t>(count));
    for (std::int32_t s = 0; s < count; ++s)
    {
        for (std::int32_t p = fFirst[s]; p < fFirst[s + 1]; ++p)
            cols[s] += weightAt[p];
        below[s] = belowWeight[fFirst[s + 1] - 1];
        top[s] = s;
    }
    const std::int32_t lastMergeable = trailing > 0 ? count - 1 : count;
    for (std::int32_t s = count - 2; s >= 0; --s)
    {
        const std::int32_t p = fParent[s];
        if (p != s + 1 || p >= lastMergeable)
            continue;
        const std::int32_t t = top[p];
        const std::int64_t mergedColumns = cols[s] + cols[t];
        const std::int64_t mergedZeros = zeros[s] + zeros[t] + cols[s] * (cols[t] + below[t] - below[s]);
        if (!mergeIsWorthIt(thresholds, mergedColumns, mergedZeros, below[t]))
            continue;
        cols[t] = mergedColumns;
        zeros[t] = mergedZeros;
        top[s] = t;
    }
    std::vector<std::int32_t> group(static_cast<std::size_t>(count));
    std::int32_t groups = 0;
    for (std::int32_t s = 0; s < count; ++s)
    {
        if (s == 0 || top[s] != top[s - 1])
            ++groups;
        group[s] = groups - 1;
    }
    result.superFirstPos.assign(static_cast<std::size_t>(groups) + 1, 0);
    result.superParent.assign(static_cast<std::size_t>(groups), ROOT);
    for (std::int32_t s = count - 1; s >= 0; --s)
        result.superFirstPos[group[s]] = fFirst[s];
    result.superFirstPos[groups] = n;
    for (std::int32_t s = 0; s < count; ++s)
    {
        if (top[s] != s)
            continue;
        result.superParent[group[s]] = fParent[s] == ROOT ? ROOT : group[fParent[s]];
    }

    // Row structures: R(s) = (adjacency of its nodes ∪ R(children)) beyond its last position.
    const std::int32_t S = groups;
    result.rowPosStart.assign(static_cast<std::size_t>(S) + 1, 0);
    result.rowStart.assign(static_cast<std::size_t>(S) + 1, 0);
    for (std::int32_t s = 0; s < S; ++s)
    {
        const std::int32_t last = result.superFirstPos[s + 1] - 1;
        result.rowPosStart[s + 1] = result.rowPosStart[s] + belowNodes[last];
        result.rowStart[s + 1] = result.rowStart[s] + belowWeight[last];
    }
    result.rowPos.resize(static_cast<std::size_t>(result.rowPosStart[S]));
    result.rows.resize(static_cast<std::size_t>(result.rowStart[S]));
    std::vector<std::vector<std::int32_t>> kids(static_cast<std::size_t>(S));
    for (std::int32_t s = 0; s < S; ++s)
        if (result.superParent[s] != ROOT)
            kids[result.superParent[s]].push_back(s);
    std::fill(mark.begin(), mark.end(), -1);
    for (std::int32_t s = 0; s < S; ++s)
    {
        const std::int32_t last = result.superFirstPos[s + 1] - 1;
        std::int64_t cursor = result.rowPosStart[s];
        const std::int64_t end = result.rowPosStart[s + 1];
        auto add = [&](std::int32_t p)
        {
            if (p <= last || mark[p] == s)
                return;
            mark[p] = s;
            if (cursor >= end)
                throw std::logic_error("symbolic: a row structure exceeds its column count");
            result.rowPos[cursor++] = p;
        };
        for (std::int32_t p = result.superFirstPos[s]; p <= last; ++p)
        {
            const std::int32_t node = order[p];
            for (std::int64_t e = off[node]; e < off[node + 1]; ++e)
                add(pos[adj[e]]);
        }
        for (const std::int32_t c : kids[s])
            for (std::int64_t k = result.rowPosStart[c]; k < result.rowPosStart[c + 1]; ++k)
                add(result.rowPos[k]);
        if (cursor != end)
            throw std::logic_error("symbolic: a row structure differs from its column count");
        std::sort(result.rowPos.begin() + result.rowPosStart[s], result.rowPos.begin() + end);
        std::int64_t dof = result.rowStart[s];
        for (std::int64_t k = result.rowPosStart[s]; k < end; ++k)
        {
            const std::int32_t p = result.rowPos[k];
            for (std::int32_t c = 0; c < weightAt[p]; ++c)
                result.rows[dof++] = result.firstColumnOfPos[p] + c;
        }
    }

    // Sizes, flops and Liu's stack of update matrices in this postorder.
    std::vector<std::int64_t> pending(static_cast<std::size_t>(S), 0);
    std::int64_t stack = 0;
    for (std::int32_t s = 0; s < S; ++s)
    {
        const std::int64_t c = result.firstColumnOfPos[result.superFirstPos[s + 1]] -
                               result.firstColumnOfPos[result.superFirstPos[s]];
        const std::int64_t r = result.rowStart[s + 1] - result.rowStart[s];
        result.panelValueCount += c * (c + r);
        result.factorEntryCount += c * (c + 1) / 2 + c * r;
        for (std::int64_t m = r; m < r + c; ++m)
            result.flopCount += static_cast<double>(m) * static_cast<double>(m + 1) + 2.0 * static_cast<double>(m);
        const std::int64_t front = (c + r) * (c + r);
        result.largestFrontValues = std::max(result.largestFrontValues, front);
        result.updateStackPeakValues = std::max(result.updateStackPeakValues, stack + front);
        stack -= pending[s];
        const std::int32_t up = result.superParent[s];
        if (up != ROOT && r > 0)
        {
            stack += r * r;
            pending[up] += r * r;
        }
    }
    result.nodeOrder = std::move(order);
    return result;
}

thread_local std::string lastError;

} // namespace

extern "C" {

const char* oosa_last_error()
{
    return lastError.c_str();
}

// Nested dissection of the node graph on node coordinates (3 per node): out_order[p] = node at position p.
int oosa_order_dissection(std::int32_t n, const std::int64_t* offsets, const std::int32_t* adjacency,
                          const std::int32_t* weights, const double* coords, std::int32_t leaf_nodes,
                          std::int32_t* out_order)
{
    try
    {
        Dissection d{n, offsets, adjacency, weights, coords, leaf_nodes > 0 ? leaf_nodes : 64, {}, {}, {}, {}};
        d.order.reserve(static_cast<std::size_t>(n));
        d.region.assign(static_cast<std::size_t>(n), -1);
        d.side.assign(static_cast<std::size_t>(n), 0);
        d.seen.assign(static_cast<std::size_t>(n), 0);
        std::vector<std::int32_t> all(static_cast<std::size_t>(n));
        std::iota(all.begin(), all.end(), 0);
        d.dissect(std::move(all));
        if (static_cast<std::int32_t>(d.order.size()) != n)
            throw std::logic_error("dissection: the order is not a permutation");
        std::copy(d.order.begin(), d.order.end(), out_order);
        return 0;
    }
    catch (const std::exception& e)
    {
        lastError = e.what();
        return 1;
    }
}

// The symbolic analysis of the node graph under `order` (node at each position; the last `trailing` positions kept
// last). amalgamate: 1 the relaxed amalgamation, 0 the fundamental supernodes. Returns a handle or null.
void* oosa_analyse(std::int32_t n, const std::int64_t* offsets, const std::int32_t* adjacency,
                   const std::int32_t* weights, const std::int32_t* order, std::int32_t trailing, int amalgamate)
{
    try
    {
        if (trailing < 0 || trailing > n)
            throw std::invalid_argument("symbolic: the trailing count is outside [0, n]");
        std::vector<char> seen(static_cast<std::size_t>(n), 0);
        for (std::int32_t p = 0; p < n; ++p)
        {
            if (order[p] < 0 || order[p] >= n || seen[order[p]])
                throw std::invalid_argument("symbolic: the order is not a permutation of the nodes");
            seen[order[p]] = 1;
        }
        for (std::int32_t v = 0; v < n; ++v)
            if (weights[v] < 1)
                throw std::invalid_argument("symbolic: a node weight is below 1");
        return new Symbolic(analyse(n, offsets, adjacency, weights, order, trailing, amalgamate != 0));
    }
    catch (const std::exception& e)
    {
        lastError = e.what();
        return nullptr;
    }
}

void oosa_free(void* handle)
{
    delete static_cast<Symbolic*>(handle);
}

// Same layout as the tests' reference driver: {nodes, DOFs, supernodes, row total, panel values, factor entries,
// exact entries}, flops.
void oosa_sizes(void* handle, std::int64_t* sizes, double* flops)
{
    const Symbolic& s = *static_cast<Symbolic*>(handle);
    sizes[0] = s.nodeCount;
    sizes[1] = s.dofCount;
    sizes[2] = static_cast<std::int64_t>(s.superParent.size());
    sizes[3] = static_cast<std::int64_t>(s.rows.size());
    sizes[4] = s.panelValueCount;
    sizes[5] = s.factorEntryCount;
    sizes[6] = s.exactEntryCount;
    sizes[7] = s.trailingDofCount;
    *flops = s.flopCount;
}

void oosa_export(void* handle, std::int32_t* node_order, std::int32_t* first_column, std::int32_t* column_count,
                 std::int32_t* parent, std::int64_t* row_start, std::int32_t* rows)
{
    const Symbolic& s = *static_cast<Symbolic*>(handle);
    std::copy(s.nodeOrder.begin(), s.nodeOrder.end(), node_order);
    const std::size_t S = s.superParent.size();
    for (std::size_t k = 0; k < S; ++k)
    {
        first_column[k] = s.firstColumnOfPos[s.superFirstPos[k]];
        column_count[k] = s.firstColumnOfPos[s.superFirstPos[k + 1]] - first_column[k];
        parent[k] = s.superParent[k];
    }
    std::copy(s.rowStart.begin(), s.rowStart.end(), row_start);
    std::copy(s.rows.begin(), s.rows.end(), rows);
}

// {largest front values, update stack peak values}.
void oosa_memory(void* handle, std::int64_t* values)
{
    const Symbolic& s = *static_cast<Symbolic*>(handle);
    values[0] = s.largestFrontValues;
    values[1] = s.updateStackPeakValues;
}

} // extern "C"
