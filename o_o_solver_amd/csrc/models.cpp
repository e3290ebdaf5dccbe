// Synthetic finite-element models: the inputs of the factor and solve (stiffness matrices in the lower compressed
// sparse column form of the reference's SymmetricCsc, include/oos/linalg/sparse/SymmetricCsc.hpp:14-20), made on the
// host. The reference assembles its matrices from real elements (src/elements, src/assembly); those are outside the
// numeric phase, so the models here are this project's own:
//   - a flat plate of quadrilateral shells, 6 DOFs per node (u, v, w, θx, θy, θz): plane-stress membrane (2 × 2
//     Gauss), Mindlin bending with MITC4 assumed transverse shear (Bathe & Dvorkin 1985), and a drilling penalty
//     γ·∫(θz − ω)² dA tying θz to the in-plane rotation ω = ½(∂v/∂x − ∂u/∂y) (Hughes & Brezzi 1989);
//   - a block of 8-node hexahedra, 3 DOFs per node (trilinear, 2 × 2 × 2 Gauss, isotropic);
//   - a stiffened cylinder: the flat shells on the cylinder's facets (each quad's stiffness rotated from its own
//     plane), bar stringers along meridians, and the geometric stiffness K_G of a uniform axial membrane force, the
//     buckling pencil K + λ·K_G (K_G on K's pattern).
// Fixed nodes (clamped edges, a fixed face) are removed with all their DOFs: every node kept has the same weight.
// The matrix is stored by 6 × 6 (or 3 × 3) node blocks over the node graph, in both directions, then written as the
// lower CSC of the DOFs in a given node order (the factor's), structural zeros kept as the reference keeps them.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace
{

// splitmix64: a small, portable generator, so that a seed gives the same model on every machine.
struct Random
{
    std::uint64_t state;
    explicit Random(std::uint64_t seed) : state(seed) {}
    std::uint64_t next()
    {
        std::uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); } // [0, 1)
    double normal()
    {
        // Box-Muller, one value per call (the second is dropped: simplicity over speed).
        double u1 = uniform();
        while (u1 <= 0.0)
            u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Model
{
    int dofPerNode = 6;
    std::int32_t nodeCount = 0; // kept nodes
    std::vector<double> coords; // 3 per kept node
    // Node graph of the kept nodes (neighbours sharing an element, self excluded), sorted.
    std::vector<std::int64_t> adjOffsets;
    std::vector<std::int32_t> adj;
    // Blocks: for kept node a, the block (a, a) then (a, adj[k]) for its neighbours, each dof² values row-major
    // (row = DOF of a, column = DOF of the neighbour). blockStart[a] = a + adjOffsets[a].
    std::vector<double> blocks;
    // Consistent load of a uniform pressure (w DOF), per kept node.
    std::vector<double> pressureLoad;
    // The geometric stiffness, in the layout of `blocks` (empty: none).
    std::vector<double> geometric;
};

std::size_t blockIndex(const Model& m, std::int32_t a, std::int32_t b)
{
    const std::int64_t base = a + m.adjOffsets[a];
    if (a == b)
        return static_cast<std::size_t>(base);
    const auto first = m.adj.begin() + m.adjOffsets[a];
    const auto last = m.adj.begin() + m.adjOffsets[a + 1];
    const auto it = std::lower_bound(first, last, b);
    if (it == last || *it != b)
        throw std::logic_error("models: block outside the node graph");
    return static_cast<std::size_t>(base + 1 + (it - first));
}

// Builds the node graph from element connectivity (kept node ids, -1 for a fixed node).
void buildGraph(Model& m, const std::vector<std::int32_t>& conn, int nodesPerElement)
{
    const std::size_t elements = conn.size() / static_cast<std::size_t>(nodesPerElement);
    std::vector<std::vector<std::int32_t>> nb(static_cast<std::size_t>(m.nodeCount));
    for (std::size_t e = 0; e < elements; ++e)
        for (int i = 0; i < nodesPerElement; ++i)
        {
            const std::int32_t a = conn[e * nodesPerElement + i];
            if (a < 0)
                continue;
            for (int j = 0; j < nodesPerElement; ++j)
            {
                const std::int32_t b = conn[e * nodesPerElement + j];
                if (b >= 0 && b != a)
                    nb[static_cast<std::size_t>(a)].push_back(b);
            }
        }
    m.adjOffsets.assign(static_cast<std::size_t>(m.nodeCount) + 1, 0);
    for (std::int32_t a = 0; a < m.nodeCount; ++a)
    {
        auto& list = nb[static_cast<std::size_t>(a)];
        std::sort(list.begin(), list.end());
        list.erase(std::unique(list.begin(), list.end()), list.end());
        m.adjOffsets[a + 1] = m.adjOffsets[a] + static_cast<std::int64_t>(list.size());
    }
    m.adj.resize(static_cast<std::size_t>(m.adjOffsets[m.nodeCount]));
    for (std::int32_t a = 0; a < m.nodeCount; ++a)
        std::copy(nb[a].begin(), nb[a].end(), m.adj.begin() + m.adjOffsets[a]);
    const std::size_t d2 = static_cast<std::size_t>(m.dofPerNode * m.dofPerNode);
    m.blocks.assign((static_cast<std::size_t>(m.nodeCount) + m.adj.size()) * d2, 0.0);
}

void scatterInto(Model& m, std::vector<double>& target, const std::int32_t* nodes, int count, const double* ke)
{
    const int d = m.dofPerNode;
    const int size = count * d;
    for (int i = 0; i < count; ++i)
    {
        const std::int32_t a = nodes[i];
        if (a < 0)
            continue;
        for (int j = 0; j < count; ++j)
        {
            const std::int32_t b = nodes[j];
            if (b < 0)
                continue;
            double* block = target.data() + blockIndex(m, a, b) * static_cast<std::size_t>(d * d);
            for (int p = 0; p < d; ++p)
                for (int q = 0; q < d; ++q)
                    block[p * d + q] += ke[(i * d + p) * size + (j * d + q)];
        }
    }
}

void scatterElement(Model& m, const std::int32_t* nodes, int count, const double* ke)
{
    scatterInto(m, m.blocks, nodes, count, ke);
}

// ---- quadrilateral flat shell ------------------------------------------------------------------------------------

const double GAUSS2[2] = {-0.5773502691896257, 0.5773502691896257};

void shapeQ4(double xi, double eta, double n[4], double dxi[4], double deta[4])
{
    const double sx[4] = {-1, 1, 1, -1};
    const double sy[4] = {-1, -1, 1, 1};
    for (int i = 0; i < 4; ++i)
    {
        n[i] = 0.25 * (1 + sx[i] * xi) * (1 + sy[i] * eta);
        dxi[i] = 0.25 * sx[i] * (1 + sy[i] * eta);
        deta[i] = 0.25 * sy[i] * (1 + sx[i] * xi);
    }
}

// ke (24 × 24, DOFs u v w θx θy θz per node) of a flat quad in the xy-plane.
void shellQuadStiffness(const double x[4], const double y[4], double E, double nu, double t, double* ke)
{
    std::fill(ke, ke + 24 * 24, 0.0);
    const double G = E / (2.0 * (1.0 + nu));
    const double dm = E * t / (1.0 - nu * nu);
    const double db = E * t * t * t / (12.0 * (1.0 - nu * nu));
    const double ds = 5.0 / 6.0 * G * t;
    const double drill = 1.0e-3 * G * t; // drilling penalty γ
    const double Cm[3][3] = {{dm, dm * nu, 0}, {dm * nu, dm, 0}, {0, 0, dm * (1 - nu) / 2}};
    const double Cb[3][3] = {{db, db * nu, 0}, {db * nu, db, 0}, {0, 0, db * (1 - nu) / 2}};
    // Index of DOF k of node i in the element vector.
    auto at = [](int i, int k) { return i * 6 + k; };

    // MITC4 tying points: γ_ξ at A = (0, 1), C = (0, −1); γ_η at B = (−1, 0), D = (1, 0).
    // Covariant shear at a point: γ_ξ = ∂w/∂ξ + βx·x_ξ + βy·y_ξ, with βx = θy, βy = −θx.
    auto covariant = [&](double xi, double eta, bool alongXi, double row[24])
    {
        double n[4], dxi[4], deta[4];
        shapeQ4(xi, eta, n, dxi, deta);
        const double* dd = alongXi ? dxi : deta;
        double gx = 0, gy = 0;
        for (int i = 0; i < 4; ++i)
        {
            gx += dd[i] * x[i];
            gy += dd[i] * y[i];
        }
        std::fill(row, row + 24, 0.0);
        for (int i = 0; i < 4; ++i)
        {
            row[at(i, 2)] = dd[i];
            row[at(i, 4)] = n[i] * gx;  // βx = θy
            row[at(i, 3)] = -n[i] * gy; // βy = −θx
        }
    };
    double tA[24], tC[24], tB[24], tD[24];
    covariant(0, 1, true, tA);
    covariant(0, -1, true, tC);
    covariant(-1, 0, false, tB);
    covariant(1, 0, false, tD);

    for (int gi = 0; gi < 2; ++gi)
        for (int gj = 0; gj < 2; ++gj)
        {
            const double xi = GAUSS2[gi], eta = GAUSS2[gj];
            double n[4], dxi[4], deta[4];
            shapeQ4(xi, eta, n, dxi, deta);
            double J[2][2] = {{0, 0}, {0, 0}};
            for (int i = 0; i < 4; ++i)
            {
                J[0][0] += dxi[i] * x[i];
                J[0][1] += dxi[i] * y[i];
                J[1][0] += deta[i] * x[i];
                J[1][1] += deta[i] * y[i];
            }
            const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
            if (det <= 0)
                throw std::runtime_error("models: an element is inverted (jitter too large)");
            const double Ji[2][2] = {{J[1][1] / det, -J[0][1] / det}, {-J[1][0] / det, J[0][0] / det}};
            double nx[4], ny[4];
            for (int i = 0; i < 4; ++i)
            {
                nx[i] = Ji[0][0] * dxi[i] + Ji[0][1] * deta[i];
                ny[i] = Ji[1][0] * dxi[i] + Ji[1][1] * deta[i];
            }
            const double w = det; // Gauss weights are 1
            // Membrane: ε = [u_x, v_y, u_y + v_x].
            double Bm[3][24] = {};
            for (int i = 0; i < 4; ++i)
            {
                Bm[0][at(i, 0)] = nx[i];
                Bm[1][at(i, 1)] = ny[i];
                Bm[2][at(i, 0)] = ny[i];
                Bm[2][at(i, 1)] = nx[i];
            }
            // Bending: κ = [βx_x, βy_y, βx_y + βy_x], βx = θy, βy = −θx.
            double Bb[3][24] = {};
            for (int i = 0; i < 4; ++i)
            {
                Bb[0][at(i, 4)] = nx[i];
                Bb[1][at(i, 3)] = -ny[i];
                Bb[2][at(i, 4)] = ny[i];
                Bb[2][at(i, 3)] = -nx[i];
            }
            // Assumed shear: covariant components interpolated from the tying points, then to Cartesian.
            double Bs[2][24];
            for (int k = 0; k < 24; ++k)
            {
                const double gxi = 0.5 * (1 + eta) * tA[k] + 0.5 * (1 - eta) * tC[k];
                const double geta = 0.5 * (1 + xi) * tD[k] + 0.5 * (1 - xi) * tB[k];
                Bs[0][k] = Ji[0][0] * gxi + Ji[0][1] * geta;
                Bs[1][k] = Ji[1][0] * gxi + Ji[1][1] * geta;
            }
            // Drilling: θz − ½(v_x − u_y).
            double Bd[24] = {};
            for (int i = 0; i < 4; ++i)
            {
                Bd[at(i, 5)] = n[i];
                Bd[at(i, 1)] = -0.5 * nx[i];
                Bd[at(i, 0)] = 0.5 * ny[i];
            }
            for (int a = 0; a < 24; ++a)
                for (int b = 0; b < 24; ++b)
                {
                    double s = 0;
                    for (int p = 0; p < 3; ++p)
                        for (int q = 0; q < 3; ++q)
                            s += Bm[p][a] * Cm[p][q] * Bm[q][b] + Bb[p][a] * Cb[p][q] * Bb[q][b];
                    s += ds * (Bs[0][a] * Bs[0][b] + Bs[1][a] * Bs[1][b]);
                    s += drill * Bd[a] * Bd[b];
                    ke[a * 24 + b] += s * w;
                }
        }
}

double quadArea(const double x[4], const double y[4])
{
    return 0.5 * std::fabs((x[0] - x[2]) * (y[1] - y[3]) - (x[1] - x[3]) * (y[0] - y[2]));
}

// ---- hexahedron ----------------------------------------------------------------------------------------------------

void hexStiffness(const double* xyz, double E, double nu, double* ke)
{
    std::fill(ke, ke + 24 * 24, 0.0);
    const double l = E * nu / ((1 + nu) * (1 - 2 * nu));
    const double mu = E / (2 * (1 + nu));
    const int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    const int sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    const int sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    for (int g = 0; g < 8; ++g)
    {
        const double xi = GAUSS2[g & 1], eta = GAUSS2[(g >> 1) & 1], zeta = GAUSS2[(g >> 2) & 1];
        double dn[8][3];
        for (int i = 0; i < 8; ++i)
        {
            dn[i][0] = 0.125 * sx[i] * (1 + sy[i] * eta) * (1 + sz[i] * zeta);
            dn[i][1] = 0.125 * sy[i] * (1 + sx[i] * xi) * (1 + sz[i] * zeta);
            dn[i][2] = 0.125 * sz[i] * (1 + sx[i] * xi) * (1 + sy[i] * eta);
        }
        double J[3][3] = {};
        for (int i = 0; i < 8; ++i)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    J[a][b] += dn[i][a] * xyz[i * 3 + b];
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        double Ji[3][3];
        Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
        Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
        Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
        Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        double d[8][3];
        for (int i = 0; i < 8; ++i)
            for (int a = 0; a < 3; ++a)
                d[i][a] = Ji[a][0] * dn[i][0] + Ji[a][1] * dn[i][1] + Ji[a][2] * dn[i][2];
        double B[6][24] = {};
        for (int i = 0; i < 8; ++i)
        {
            B[0][3 * i] = d[i][0];
            B[1][3 * i + 1] = d[i][1];
            B[2][3 * i + 2] = d[i][2];
            B[3][3 * i] = d[i][1];
            B[3][3 * i + 1] = d[i][0];
            B[4][3 * i + 1] = d[i][2];
            B[4][3 * i + 2] = d[i][1];
            B[5][3 * i] = d[i][2];
            B[5][3 * i + 2] = d[i][0];
        }
        const double C[6] = {l + 2 * mu, l, mu};
        for (int a = 0; a < 24; ++a)
            for (int b = 0; b < 24; ++b)
            {
                double s = 0;
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q)
                        s += B[p][a] * (p == q ? C[0] : C[1]) * B[q][b];
                for (int p = 3; p < 6; ++p)
                    s += B[p][a] * C[2] * B[p][b];
                ke[a * 24 + b] += s * std::fabs(det);
            }
    }
}

} // namespace

extern "C" {

// A clamped flat plate of nx × ny quadrilateral shells over [0, lx] × [0, ly]: every node of the four edges fixed.
// jitter: interior nodes moved in-plane by U(−jitter, jitter)·h per coordinate. thickness: thick_mode 0 constant
// t_lo, 1 log-uniform in [t_lo, t_hi] per element, 2 uniform in [t_lo, t_hi] per element.
void* oosm_plate(std::int32_t nx, std::int32_t ny, double lx, double ly, double jitter, std::uint64_t seed,
                 double t_lo, double t_hi, int thick_mode, double E, double nu, double pressure)
{
    auto m = std::make_unique<Model>();
    m->dofPerNode = 6;
    Random random(seed);
    const std::int32_t px = nx + 1, py = ny + 1;
    const double hx = lx / nx, hy = ly / ny;
    std::vector<double> X(static_cast<std::size_t>(px) * py), Y(X.size());
    std::vector<std::int32_t> kept(X.size(), -1);
    std::int32_t count = 0;
    for (std::int32_t j = 0; j < py; ++j)
        for (std::int32_t i = 0; i < px; ++i)
        {
            const std::size_t g = static_cast<std::size_t>(j) * px + i;
            const bool interior = i > 0 && j > 0 && i < nx && j < ny;
            double x = i * hx, y = j * hy;
            if (interior && jitter > 0)
            {
                x += (2 * random.uniform() - 1) * jitter * hx;
                y += (2 * random.uniform() - 1) * jitter * hy;
            }
            X[g] = x;
            Y[g] = y;
            if (interior)
                kept[g] = count++;
        }
    m->nodeCount = count;
    m->coords.resize(static_cast<std::size_t>(count) * 3);
    for (std::size_t g = 0; g < X.size(); ++g)
        if (kept[g] >= 0)
        {
            m->coords[3 * kept[g]] = X[g];
            m->coords[3 * kept[g] + 1] = Y[g];
            m->coords[3 * kept[g] + 2] = 0.0;
        }
    std::vector<std::int32_t> conn(static_cast<std::size_t>(nx) * ny * 4);
    for (std::int32_t j = 0; j < ny; ++j)
        for (std::int32_t i = 0; i < nx; ++i)
        {
            const std::size_t e = static_cast<std::size_t>(j) * nx + i;
            const std::size_t g0 = static_cast<std::size_t>(j) * px + i;
            conn[4 * e] = kept[g0];
            conn[4 * e + 1] = kept[g0 + 1];
            conn[4 * e + 2] = kept[g0 + 1 + px];
            conn[4 * e + 3] = kept[g0 + px];
        }
    buildGraph(*m, conn, 4);
    m->pressureLoad.assign(static_cast<std::size_t>(count), 0.0);
    std::vector<double> ke(24 * 24);
    for (std::int32_t j = 0; j < ny; ++j)
        for (std::int32_t i = 0; i < nx; ++i)
        {
            const std::size_t e = static_cast<std::size_t>(j) * nx + i;
            const std::size_t g0 = static_cast<std::size_t>(j) * px + i;
            const std::size_t g[4] = {g0, g0 + 1, g0 + 1 + px, g0 + px};
            double x[4], y[4];
            for (int k = 0; k < 4; ++k)
            {
                x[k] = X[g[k]];
                y[k] = Y[g[k]];
            }
            double t = t_lo;
            if (thick_mode == 1)
                t = std::exp(std::log(t_lo) + random.uniform() * (std::log(t_hi) - std::log(t_lo)));
            else if (thick_mode == 2)
                t = t_lo + random.uniform() * (t_hi - t_lo);
            shellQuadStiffness(x, y, E, nu, t, ke.data());
            scatterElement(*m, &conn[4 * e], 4, ke.data());
            const double share = pressure * quadArea(x, y) / 4.0;
            for (int k = 0; k < 4; ++k)
                if (conn[4 * e + k] >= 0)
                    m->pressureLoad[static_cast<std::size_t>(conn[4 * e + k])] += share;
        }
    return m.release();
}

// A block of nx × ny × nz hexahedra over [0, lx] × [0, ly] × [0, lz], the nodes of the face z = 0 fixed; per-element
// Young's modulus E_median·exp(sigma·N(0, 1)) (sigma = 0: uniform).
void* oosm_hex_block(std::int32_t nx, std::int32_t ny, std::int32_t nz, double lx, double ly, double lz,
                     std::uint64_t seed, double E_median, double sigma, double nu)
{
    auto m = std::make_unique<Model>();
    m->dofPerNode = 3;
    Random random(seed);
    const std::int32_t px = nx + 1, py = ny + 1, pz = nz + 1;
    const std::size_t total = static_cast<std::size_t>(px) * py * pz;
    std::vector<std::int32_t> kept(total, -1);
    std::int32_t count = 0;
    for (std::int32_t k = 0; k < pz; ++k)
        for (std::int32_t j = 0; j < py; ++j)
            for (std::int32_t i = 0; i < px; ++i)
                if (k > 0)
                    kept[(static_cast<std::size_t>(k) * py + j) * px + i] = count++;
    m->nodeCount = count;
    m->coords.resize(static_cast<std::size_t>(count) * 3);
    auto point = [&](std::int32_t i, std::int32_t j, std::int32_t k, double* p)
    {
        p[0] = lx * i / nx;
        p[1] = ly * j / ny;
        p[2] = lz * k / nz;
    };
    for (std::int32_t k = 0; k < pz; ++k)
        for (std::int32_t j = 0; j < py; ++j)
            for (std::int32_t i = 0; i < px; ++i)
            {
                const std::int32_t id = kept[(static_cast<std::size_t>(k) * py + j) * px + i];
                if (id >= 0)
                    point(i, j, k, &m->coords[3 * static_cast<std::size_t>(id)]);
            }
    const std::size_t elements = static_cast<std::size_t>(nx) * ny * nz;
    std::vector<std::int32_t> conn(elements * 8);
    const int oi[8] = {0, 1, 1, 0, 0, 1, 1, 0}, oj[8] = {0, 0, 1, 1, 0, 0, 1, 1}, ok[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    for (std::int32_t k = 0; k < nz; ++k)
        for (std::int32_t j = 0; j < ny; ++j)
            for (std::int32_t i = 0; i < nx; ++i)
            {
                const std::size_t e = (static_cast<std::size_t>(k) * ny + j) * nx + i;
                for (int c = 0; c < 8; ++c)
                    conn[8 * e + c] = kept[(static_cast<std::size_t>(k + ok[c]) * py + (j + oj[c])) * px + (i + oi[c])];
            }
    buildGraph(*m, conn, 8);
    m->pressureLoad.assign(static_cast<std::size_t>(count), 0.0);
    std::vector<double> ke(24 * 24);
    double xyz[24];
    for (std::int32_t k = 0; k < nz; ++k)
        for (std::int32_t j = 0; j < ny; ++j)
            for (std::int32_t i = 0; i < nx; ++i)
            {
                const std::size_t e = (static_cast<std::size_t>(k) * ny + j) * nx + i;
                for (int c = 0; c < 8; ++c)
                    point(i + oi[c], j + oj[c], k + ok[c], &xyz[3 * c]);
                const double E = sigma > 0 ? E_median * std::exp(sigma * random.normal()) : E_median;
                hexStiffness(xyz, E, nu, ke.data());
                scatterElement(*m, &conn[8 * e], 8, ke.data());
            }
    return m.release();
}

// A stiffened cylinder of radius R and length L: n_circ × n_axial quadrilateral shells of thickness t on the facets
// of the cylinder (R cos θ, R sin θ, z), z ∈ [0, L], the nodes of both end rings fixed; bar stringers of area
// stringer_area (modulus E) along every stringer_every-th meridian; and the geometric stiffness of a uniform axial
// membrane force N (< 0: compression; force per unit length of the circumference), which the stringers carry as
// P = N·A/t: K_G = N·∫ ∂N_a/∂z ∂N_b/∂z dA on each translation of the skin, P/h·[1 −1; −1 1] on the two translations
// across a bar. The buckling factors λ solve (K + λ·K_G)φ = 0. Each facet is flat (its nodes lie on two parallel
// generators): its stiffness is formed in its own frame (x along the chord, y along z) and rotated.
void* oosm_cylinder(std::int32_t n_circ, std::int32_t n_axial, double radius, double length, double t, double E,
                    double nu, std::int32_t stringer_every, double stringer_area, double axial_force)
{
    auto m = std::make_unique<Model>();
    m->dofPerNode = 6;
    const std::int32_t rings = n_axial + 1;
    const auto node = [n_circ](std::int32_t i, std::int32_t j) { return static_cast<std::size_t>(j) * n_circ + i; };
    std::vector<std::int32_t> kept(static_cast<std::size_t>(n_circ) * rings, -1);
    std::int32_t count = 0;
    for (std::int32_t j = 1; j < n_axial; ++j)
        for (std::int32_t i = 0; i < n_circ; ++i)
            kept[node(i, j)] = count++;
    const double pi2 = 6.283185307179586;
    const auto position = [&](std::int32_t i, std::int32_t j, double* p)
    {
        const double th = pi2 * (i % n_circ) / n_circ;
        p[0] = radius * std::cos(th);
        p[1] = radius * std::sin(th);
        p[2] = length * j / n_axial;
    };
    m->nodeCount = count;
    m->coords.resize(static_cast<std::size_t>(count) * 3);
    for (std::int32_t j = 0; j < rings; ++j)
        for (std::int32_t i = 0; i < n_circ; ++i)
            if (kept[node(i, j)] >= 0)
                position(i, j, &m->coords[3 * static_cast<std::size_t>(kept[node(i, j)])]);
    std::vector<std::int32_t> conn(static_cast<std::size_t>(n_circ) * n_axial * 4);
    for (std::int32_t j = 0; j < n_axial; ++j)
        for (std::int32_t i = 0; i < n_circ; ++i)
        {
            const std::size_t e = static_cast<std::size_t>(j) * n_circ + i;
            const std::int32_t i1 = (i + 1) % n_circ;
            conn[4 * e] = kept[node(i, j)];
            conn[4 * e + 1] = kept[node(i1, j)];
            conn[4 * e + 2] = kept[node(i1, j + 1)];
            conn[4 * e + 3] = kept[node(i, j + 1)];
        }
    buildGraph(*m, conn, 4);
    m->geometric.assign(m->blocks.size(), 0.0);
    m->pressureLoad.assign(static_cast<std::size_t>(count), 0.0);
    std::vector<double> kl(24 * 24), kg(24 * 24), geo(24 * 24);
    const double h = length / n_axial;
    for (std::int32_t j = 0; j < n_axial; ++j)
        for (std::int32_t i = 0; i < n_circ; ++i)
        {
            const std::size_t e = static_cast<std::size_t>(j) * n_circ + i;
            double p0[3], p1[3];
            position(i, j, p0);
            position(i + 1, j, p1);
            double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], 0.0};
            const double chord = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1]);
            e1[0] /= chord;
            e1[1] /= chord;
            const double e2[3] = {0.0, 0.0, 1.0};
            const double e3[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                  e1[0] * e2[1] - e1[1] * e2[0]};
            const double R[3][3] = {{e1[0], e1[1], e1[2]}, {e2[0], e2[1], e2[2]}, {e3[0], e3[1], e3[2]}};
            const double x[4] = {0.0, chord, chord, 0.0}, y[4] = {0.0, 0.0, h, h};
            shellQuadStiffness(x, y, E, nu, t, kl.data());
            // kg = Tᵀ·kl·T, T the block diagonal of R over the 8 triads (translations and rotations of 4 nodes).
            for (int a = 0; a < 24; ++a)
                for (int b = 0; b < 24; ++b)
                {
                    // (kl·T)(a, b) = Σ_c kl(a, 3·(b/3) + c)·R(c, b mod 3)
                    double v = 0.0;
                    for (int c = 0; c < 3; ++c)
                        v += kl[a * 24 + 3 * (b / 3) + c] * R[c][b % 3];
                    geo[a * 24 + b] = v;
                }
            for (int a = 0; a < 24; ++a)
                for (int b = 0; b < 24; ++b)
                {
                    double v = 0.0;
                    for (int c = 0; c < 3; ++c)
                        v += R[c][a % 3] * geo[(3 * (a / 3) + c) * 24 + b];
                    kg[a * 24 + b] = v;
                }
            scatterElement(*m, &conn[4 * e], 4, kg.data());
            // The membrane force's geometric stiffness, the same on each translation (and so in any frame).
            std::fill(geo.begin(), geo.end(), 0.0);
            for (int gy = 0; gy < 2; ++gy)
                for (int gx = 0; gx < 2; ++gx)
                {
                    double n[4], dxi[4], deta[4];
                    shapeQ4(GAUSS2[gx], GAUSS2[gy], n, dxi, deta);
                    const double weight = 0.25 * chord * h; // dA = (chord/2)(h/2) dξ dη, unit Gauss weights
                    for (int a = 0; a < 4; ++a)
                        for (int b = 0; b < 4; ++b)
                        {
                            const double v = axial_force * (deta[a] * 2.0 / h) * (deta[b] * 2.0 / h) * weight;
                            for (int c = 0; c < 3; ++c)
                                geo[(6 * a + c) * 24 + 6 * b + c] += v;
                        }
                }
            scatterInto(*m, m->geometric, &conn[4 * e], 4, geo.data());
        }
    if (stringer_every > 0 && stringer_area > 0.0)
    {
        const double k = E * stringer_area / h, pk = axial_force * stringer_area / t / h;
        std::vector<double> kb(12 * 12), gb(12 * 12);
        for (std::int32_t i = 0; i < n_circ; i += stringer_every)
            for (std::int32_t j = 0; j < n_axial; ++j)
            {
                const std::int32_t bar[2] = {kept[node(i, j)], kept[node(i, j + 1)]};
                std::fill(kb.begin(), kb.end(), 0.0);
                std::fill(gb.begin(), gb.end(), 0.0);
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b)
                    {
                        const double sign = a == b ? 1.0 : -1.0;
                        kb[(6 * a + 2) * 12 + 6 * b + 2] = sign * k;
                        for (int c = 0; c < 2; ++c)
                            gb[(6 * a + c) * 12 + 6 * b + c] = sign * pk;
                    }
                scatterElement(*m, bar, 2, kb.data());
                scatterInto(*m, m->geometric, bar, 2, gb.data());
            }
    }
    return m.release();
}

void oosm_free(void* handle)
{
    delete static_cast<Model*>(handle);
}

// sizes = {kept nodes, DOFs per node, adjacency entries, lower CSC entries of the DOFs}.
void oosm_sizes(void* handle, std::int64_t* sizes)
{
    const Model& m = *static_cast<Model*>(handle);
    const std::int64_t d = m.dofPerNode;
    sizes[0] = m.nodeCount;
    sizes[1] = d;
    sizes[2] = static_cast<std::int64_t>(m.adj.size());
    // Per node: its diagonal block's lower triangle and one full block per edge.
    sizes[3] = m.nodeCount * (d * (d + 1) / 2) + static_cast<std::int64_t>(m.adj.size() / 2) * d * d;
}

void oosm_graph(void* handle, std::int64_t* offsets, std::int32_t* adjacency)
{
    const Model& m = *static_cast<Model*>(handle);
    std::copy(m.adjOffsets.begin(), m.adjOffsets.end(), offsets);
    std::copy(m.adj.begin(), m.adj.end(), adjacency);
}

void oosm_coords(void* handle, double* xyz)
{
    const Model& m = *static_cast<Model*>(handle);
    std::copy(m.coords.begin(), m.coords.end(), xyz);
}

// The lower CSC of the DOFs in the order of `node_order` (node_order[p] = node at position p; the DOFs of a node
// consecutive): offsets[n + 1], rows[nnz], values[nnz], each column's rows ascending from its diagonal.
void writeCsc(const Model& m, const std::vector<double>& source, const std::int32_t* node_order,
              std::int64_t* offsets, std::int32_t* rows, double* values);

void oosm_csc(void* handle, const std::int32_t* node_order, std::int64_t* offsets, std::int32_t* rows,
              double* values)
{
    const Model& m = *static_cast<Model*>(handle);
    writeCsc(m, m.blocks, node_order, offsets, rows, values);
}

// The geometric stiffness on the same pattern (a model without one: all zero).
void oosm_csc_geometric(void* handle, const std::int32_t* node_order, std::int64_t* offsets, std::int32_t* rows,
                        double* values)
{
    const Model& m = *static_cast<Model*>(handle);
    if (m.geometric.empty())
    {
        std::vector<double> zero(m.blocks.size(), 0.0);
        writeCsc(m, zero, node_order, offsets, rows, values);
    }
    else
        writeCsc(m, m.geometric, node_order, offsets, rows, values);
}

} // extern "C"

void writeCsc(const Model& m, const std::vector<double>& source, const std::int32_t* node_order,
              std::int64_t* offsets, std::int32_t* rows, double* values)
{
    const int d = m.dofPerNode;
    const std::int32_t n = m.nodeCount;
    std::vector<std::int32_t> position(static_cast<std::size_t>(n));
    for (std::int32_t p = 0; p < n; ++p)
        position[static_cast<std::size_t>(node_order[p])] = p;
    std::int64_t at = 0;
    offsets[0] = 0;
    std::vector<std::pair<std::int32_t, std::int32_t>> later; // (position, node) of the neighbours after p
    for (std::int32_t p = 0; p < n; ++p)
    {
        const std::int32_t a = node_order[p];
        later.clear();
        for (std::int64_t k = m.adjOffsets[a]; k < m.adjOffsets[a + 1]; ++k)
        {
            const std::int32_t b = m.adj[static_cast<std::size_t>(k)];
            if (position[static_cast<std::size_t>(b)] > p)
                later.emplace_back(position[static_cast<std::size_t>(b)], b);
        }
        std::sort(later.begin(), later.end());
        const double* diag = source.data() + blockIndex(m, a, a) * static_cast<std::size_t>(d * d);
        for (int q = 0; q < d; ++q)
        {
            // Column q of node a: the diagonal block's rows q..d−1, then each later node's d rows.
            for (int r = q; r < d; ++r)
            {
                rows[at] = p * d + r;
                values[at] = diag[r * d + q];
                ++at;
            }
            for (const auto& [pb, b] : later)
            {
                // Entry (DOF r of b, DOF q of a) = block(b, a)[r][q].
                const double* block = source.data() + blockIndex(m, b, a) * static_cast<std::size_t>(d * d);
                for (int r = 0; r < d; ++r)
                {
                    rows[at] = pb * d + r;
                    values[at] = block[r * d + q];
                    ++at;
                }
            }
            offsets[p * d + q + 1] = at;
        }
    }
}

extern "C" {

// Loads in the DOF order of `node_order`, column-major n × nrhs. kind 0: the uniform pressure (w DOF, one column,
// shell models); kind 1: independent N(0, 1) forces on the translational DOFs (3 per node), rotations 0; kind 2:
// N(0, 1) forces on the translational DOFs of the nodes whose coordinate z is the largest (a top face), others 0.
void oosm_loads(void* handle, const std::int32_t* node_order, int kind, std::int32_t nrhs, std::uint64_t seed,
                double* b)
{
    const Model& m = *static_cast<Model*>(handle);
    const int d = m.dofPerNode;
    const std::int64_t n = static_cast<std::int64_t>(m.nodeCount) * d;
    std::fill(b, b + n * nrhs, 0.0);
    Random random(seed);
    double top = -1e300;
    for (std::int32_t a = 0; a < m.nodeCount; ++a)
        top = std::max(top, m.coords[3 * static_cast<std::size_t>(a) + 2]);
    // Drawn in node order, so that the loads do not depend on the ordering.
    std::vector<std::int32_t> position(static_cast<std::size_t>(m.nodeCount));
    for (std::int32_t p = 0; p < m.nodeCount; ++p)
        position[static_cast<std::size_t>(node_order[p])] = p;
    for (std::int32_t c = 0; c < nrhs; ++c)
        for (std::int32_t a = 0; a < m.nodeCount; ++a)
        {
            const std::int64_t p = position[static_cast<std::size_t>(a)];
            double* at = b + c * n + p * d;
            if (kind == 0)
                at[2] = m.pressureLoad[static_cast<std::size_t>(a)];
            else if (kind == 1 || (kind == 2 && m.coords[3 * static_cast<std::size_t>(a) + 2] >= top))
                for (int k = 0; k < 3; ++k)
                    at[k] = random.normal();
        }
}

} // extern "C"
