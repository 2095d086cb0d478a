// Input formats on the host: Gmsh MSH2 structured grids, .silpro problem
// files, and the cell inputs of the magnetostatics preprocess.
//
// * MSH2 (onelab_interface/gmsh_structured_grid.hpp:239-553): ASCII version
//   2 only; a mesh made entirely of quadrilaterals (type 3) or entirely of
//   hexahedra (type 5), with points, lines and triangles (types 15, 1, 2)
//   skipped as boundary elements;
//   the nodes must form a full rectilinear grid (coordinates merged within
//   1e-12) and every cell must be exactly one grid cell.  The reference
//   reads through iostreams, keeps node tags in a std::map and locates every
//   coordinate by a linear scan of the axis; here the file is read in one
//   block and tokenised in place with std::from_chars, axes are located by
//   binary search (keeping the reference's rule: the first axis value within
//   the tolerance), and node tags index a dense table when they are dense.
// * .silpro (the format of onelab_interface/onelab_interface.hpp:160-560):
//   read by a scannerless recursive descent straight from the text (grammar
//   beside SilproReader) into a store keyed by block path; the first of a
//   repeated key or block wins; the solver settings with the defaults of
//   MinimizeStrongFormulationResidualProblem (:96-111).
// * Cell inputs (magnetostatics_onelab.hpp:4870-4897): μ, current density
//   and the nonlinear-material flag of every cell from its physical tag.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <new>
#include <set>
#include <string_view>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "common.hpp"

struct sil_structured_grid_s {
    int ndim = 0;
    std::vector<double> coords[3];
    double z_value = 0.0;
    std::vector<int32_t> cell_tags; // x fastest
};

// A parsed .silpro file: every assignment under the path of its key (the
// names of the enclosing blocks, then the key), and the path of every block.
// Paths compare element by element, so the assignments of one block are
// contiguous and in key order.
struct sil_silpro_s {
    using Path = std::vector<std::string>;
    std::map<Path, std::string> assignments;
    std::set<Path> blocks;
};

namespace sil {
namespace {

constexpr double kCoordTol = 1.0e-12;

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ------------------------------------------------------------------ MSH2 ---

class Tokens {
  public:
    Tokens(const char* b, const char* e) : p_(b), e_(e) {}
    bool next(std::string_view& t)
    {
        while (p_ < e_ && is_space(*p_)) ++p_;
        if (p_ >= e_) return false;
        const char* s = p_;
        while (p_ < e_ && !is_space(*p_)) ++p_;
        t = std::string_view(s, static_cast<size_t>(p_ - s));
        return true;
    }
    const char* pos() const { return p_; }
    void seek(const char* p) { p_ = p; }
    template <typename T>
    T number()
    {
        std::string_view t;
        if (!next(t)) throw ParseError("unexpected end of mesh file");
        if (t.size() > 1 && t[0] == '+') t.remove_prefix(1); // accepted by iostream extraction
        T v{};
        auto r = std::from_chars(t.data(), t.data() + t.size(), v);
        if (r.ec != std::errc() || r.ptr != t.data() + t.size())
            throw ParseError("invalid number '" + std::string(t) + "' in mesh file");
        return v;
    }

  private:
    static bool is_space(char c) { return c == ' ' || c == '\n' || c == '\r' || c == '\t' || c == '\v' || c == '\f'; }
    const char* p_;
    const char* e_;
};

int element_dimension(int type)
{
    switch (type) {
    case 15: return 0;
    case 1: return 1;
    case 2:
    case 3: return 2;
    case 4:
    case 5:
    case 6:
    case 7: return 3;
    default: return -1;
    }
}

int nodes_per_element(int type)
{
    switch (type) {
    case 1: return 2;
    case 2: return 3;
    case 3: return 4;
    case 4: return 4;
    case 5: return 8;
    case 6: return 6;
    case 7: return 5;
    case 15: return 1;
    default: return 0;
    }
}

bool near(double a, double b) { return std::abs(a - b) < kCoordTol; }

// sorted values merged within the tolerance (consecutive values closer than
// it to the last kept one are dropped); exact repeats are removed first
// through a hash set — a grid axis repeats each value nx·ny times — which
// leaves the merge unchanged
std::vector<double> unique_sorted(const std::vector<double>& v)
{
    std::unordered_set<double> distinct(v.begin(), v.end());
    std::vector<double> s(distinct.begin(), distinct.end());
    std::sort(s.begin(), s.end());
    std::vector<double> out;
    out.reserve(s.size());
    for (double x : s)
        if (out.empty() || !near(out.back(), x)) out.push_back(x);
    return out;
}

// first axis index whose value is within the tolerance of x
size_t axis_index(const std::vector<double>& axis, double x)
{
    auto it = std::lower_bound(axis.begin(), axis.end(), x - kCoordTol);
    size_t i = static_cast<size_t>(it - axis.begin());
    // the first within tolerance is at i or i + 1 (a value below x − tol is not)
    for (size_t k = (i > 0 ? i - 1 : 0); k < std::min(axis.size(), i + 2); ++k)
        if (near(axis[k], x)) return k;
    throw ParseError("coordinate does not belong to the detected structured grid");
}

struct Mesh {
    std::vector<int64_t> node_tag;
    std::vector<double> x, y, z;
    int cell_dim = 0;              // 2 (quadrilaterals) or 3 (hexahedra)
    std::vector<int32_t> cell_tag; // physical tag per cell
    std::vector<int64_t> cell_nodes;
};

// ----- parallel reading of the node and element sections -----
//
// Gmsh writes one node or element per line; the sections are split into
// line-aligned chunks read by one thread each and merged in order.  Anything
// irregular (a record over several lines, a count that does not match, an
// element the fast path does not handle) makes the caller read the section
// serially instead, so results and errors are those of the serial reader.

int reader_threads()
{
    const unsigned hw = std::thread::hardware_concurrency();
    return static_cast<int>(std::max(1u, std::min(hw == 0 ? 1u : hw, 16u)));
}

// chunk starts: line starts of [b, e), at most T + 1 entries ending with e
std::vector<const char*> line_chunks(const char* b, const char* e, int T)
{
    std::vector<const char*> cuts{b};
    const size_t len = static_cast<size_t>(e - b);
    for (int t = 1; t < T; ++t) {
        const char* p = b + len * t / T;
        if (p <= cuts.back()) continue;
        const void* nl = std::memchr(p, '\n', static_cast<size_t>(e - p));
        if (!nl) break;
        p = static_cast<const char*>(nl) + 1;
        if (p > cuts.back() && p < e) cuts.push_back(p);
    }
    cuts.push_back(e);
    return cuts;
}

template <class F>
void run_threads(int n, F f)
{
    std::vector<std::thread> th;
    th.reserve(n);
    for (int t = 0; t < n; ++t) th.emplace_back(f, t);
    for (auto& x : th) x.join();
}

// number of non-blank lines of [b, e)
int64_t count_records(const char* b, const char* e)
{
    int64_t n = 0;
    bool content = false;
    for (const char* p = b; p < e; ++p) {
        if (*p == '\n') {
            n += content;
            content = false;
        } else if (!(*p == ' ' || *p == '\r' || *p == '\t')) {
            content = true;
        }
    }
    return n + content;
}

const char* find_marker(const char* b, const char* e, const char* marker)
{
    const size_t ml = std::strlen(marker);
    for (const char* p = b; p + ml <= e; ++p) {
        p = static_cast<const char*>(std::memchr(p, marker[0], static_cast<size_t>(e - p)));
        if (!p || p + ml > e) return nullptr;
        if (std::memcmp(p, marker, ml) == 0) return p;
    }
    return nullptr;
}

// the $Nodes body [b, end) holding `count` records; false: read serially
bool fast_nodes(const char* b, const char* end, int64_t count, Mesh& m)
{
    const int T = reader_threads();
    const auto cuts = line_chunks(b, end, T);
    const int nc = static_cast<int>(cuts.size()) - 1;
    std::vector<int64_t> n(nc + 1, 0);
    run_threads(nc, [&](int t) { n[t + 1] = count_records(cuts[t], cuts[t + 1]); });
    for (int t = 0; t < nc; ++t) n[t + 1] += n[t];
    if (n[nc] != count) return false;
    m.node_tag.resize(count);
    m.x.resize(count);
    m.y.resize(count);
    m.z.resize(count);
    std::vector<char> ok(nc, 1);
    run_threads(nc, [&](int t) {
        try {
            const char* p = cuts[t];
            int64_t r = n[t];
            while (p < cuts[t + 1]) {
                const void* nl = std::memchr(p, '\n', static_cast<size_t>(cuts[t + 1] - p));
                const char* le = nl ? static_cast<const char*>(nl) : cuts[t + 1];
                Tokens line(p, le);
                std::string_view tok;
                Tokens probe(p, le);
                if (probe.next(tok)) {
                    m.node_tag[r] = line.number<int64_t>();
                    m.x[r] = line.number<double>();
                    m.y[r] = line.number<double>();
                    m.z[r] = line.number<double>();
                    if (line.next(tok)) throw ParseError("extra token");
                    ++r;
                }
                p = le + 1;
            }
        } catch (const std::exception&) {
            ok[t] = 0;
        }
    });
    return std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; });
}

struct ElementChunk {
    std::vector<int32_t> tags;
    std::vector<int64_t> nodes;
    int dim = 0;
    bool ok = true;
};

// the $Elements body: quadrilaterals or hexahedra plus points, lines and
// triangles, one per line; false: read serially
bool fast_elements(const char* b, const char* end, int64_t count, Mesh& m)
{
    const int T = reader_threads();
    const auto cuts = line_chunks(b, end, T);
    const int nc = static_cast<int>(cuts.size()) - 1;
    std::vector<int64_t> n(nc, 0);
    run_threads(nc, [&](int t) { n[t] = count_records(cuts[t], cuts[t + 1]); });
    int64_t total = 0;
    for (int64_t v : n) total += v;
    if (total != count) return false;
    std::vector<ElementChunk> ch(nc);
    run_threads(nc, [&](int t) {
        ElementChunk& c = ch[t];
        try {
            const char* p = cuts[t];
            while (p < cuts[t + 1]) {
                const void* nl = std::memchr(p, '\n', static_cast<size_t>(cuts[t + 1] - p));
                const char* le = nl ? static_cast<const char*>(nl) : cuts[t + 1];
                Tokens line(p, le);
                std::string_view tok;
                Tokens probe(p, le);
                if (probe.next(tok)) {
                    (void)line.number<int64_t>();
                    const int type = line.number<int>();
                    const int ntags = line.number<int>();
                    int32_t physical = 0;
                    for (int q = 0; q < ntags; ++q) {
                        const int32_t v = line.number<int32_t>();
                        if (q == 0) physical = v;
                    }
                    const int nn = nodes_per_element(type);
                    if (type == 3 || type == 5) {
                        const int dim = type == 3 ? 2 : 3;
                        if (c.dim == 0) c.dim = dim;
                        if (c.dim != dim) throw ParseError("mixed");
                        c.tags.push_back(physical);
                        for (int q = 0; q < nn; ++q) c.nodes.push_back(line.number<int64_t>());
                    } else if (type == 1 || type == 2 || type == 15) {
                        for (int q = 0; q < nn; ++q) (void)line.number<int64_t>();
                    } else {
                        throw ParseError("other");
                    }
                    if (line.next(tok)) throw ParseError("extra token");
                }
                p = le + 1;
            }
        } catch (const std::exception&) {
            c.ok = false;
        }
    });
    int dim = 0;
    size_t ncell = 0;
    for (const auto& c : ch) {
        if (!c.ok) return false;
        if (c.dim != 0) {
            if (dim == 0) dim = c.dim;
            if (dim != c.dim) return false;
        }
        ncell += c.tags.size();
    }
    if (m.cell_dim != 0 && dim != 0 && dim != m.cell_dim) return false;
    if (dim != 0) m.cell_dim = dim;
    const int per = dim == 3 ? 8 : 4;
    size_t at = m.cell_tag.size();
    m.cell_tag.resize(at + ncell);
    m.cell_nodes.resize((at + ncell) * per);
    for (const auto& c : ch) {
        std::copy(c.tags.begin(), c.tags.end(), m.cell_tag.begin() + at);
        std::copy(c.nodes.begin(), c.nodes.end(), m.cell_nodes.begin() + at * per);
        at += c.tags.size();
    }
    return true;
}

Mesh parse_msh2(const char* b, const char* e)
{
    Tokens tk(b, e);
    Mesh m;
    std::string_view t;
    bool saw_elements = false;
    while (tk.next(t)) {
        if (t == "$MeshFormat") {
            std::string_view v;
            if (!tk.next(v)) throw ParseError("unexpected end of mesh file");
            double version = 0.0;
            auto r = std::from_chars(v.data(), v.data() + v.size(), version);
            if (r.ec != std::errc()) throw ParseError("invalid mesh format version");
            const int file_type = tk.number<int>();
            (void)tk.number<int>(); // data size
            if (version < 2.0 || version >= 3.0 || file_type != 0)
                throw ParseError("only Gmsh msh2 ASCII meshes are supported");
        } else if (t == "$Nodes") {
            const int64_t count = tk.number<int64_t>();
            const char* body = tk.pos();
            const char* stop = find_marker(body, e, "$EndNodes");
            if (stop && count > 4096 && fast_nodes(body, stop, count, m)) {
                tk.seek(stop);
                continue;
            }
            m.node_tag.resize(count);
            m.x.resize(count);
            m.y.resize(count);
            m.z.resize(count);
            for (int64_t i = 0; i < count; ++i) {
                m.node_tag[i] = tk.number<int64_t>();
                m.x[i] = tk.number<double>();
                m.y[i] = tk.number<double>();
                m.z[i] = tk.number<double>();
            }
        } else if (t == "$Elements") {
            saw_elements = true;
            const int64_t count = tk.number<int64_t>();
            const char* body = tk.pos();
            const char* stop = find_marker(body, e, "$EndElements");
            if (stop && count > 4096) {
                Mesh trial;
                trial.cell_dim = m.cell_dim;
                if (fast_elements(body, stop, count, trial)) {
                    if (m.cell_tag.empty()) {
                        m.cell_dim = trial.cell_dim;
                        m.cell_tag = std::move(trial.cell_tag);
                        m.cell_nodes = std::move(trial.cell_nodes);
                    } else {
                        m.cell_dim = trial.cell_dim;
                        m.cell_tag.insert(m.cell_tag.end(), trial.cell_tag.begin(), trial.cell_tag.end());
                        m.cell_nodes.insert(m.cell_nodes.end(), trial.cell_nodes.begin(), trial.cell_nodes.end());
                    }
                    tk.seek(stop);
                    continue;
                }
            }
            for (int64_t i = 0; i < count; ++i) {
                (void)tk.number<int64_t>(); // element tag
                const int type = tk.number<int>();
                const int ntags = tk.number<int>();
                int32_t physical = 0;
                for (int q = 0; q < ntags; ++q) {
                    const int32_t v = tk.number<int32_t>();
                    if (q == 0) physical = v;
                }
                if (element_dimension(type) < 0)
                    throw ParseError("unsupported Gmsh element type " + std::to_string(type) + " found in mesh");
                const int nn = nodes_per_element(type);
                if (type == 3 || type == 5) {
                    // every quadrilateral is a cell: a mesh mixing them with
                    // hexahedra is rejected, as the reference does
                    const int dim = type == 3 ? 2 : 3;
                    if (m.cell_dim == 0) m.cell_dim = dim;
                    if (m.cell_dim != dim)
                        throw ParseError("unsupported mesh topology: SimiLie currently requires the whole mesh to be "
                                         "made of quadrilaterals or hexahedra");
                    m.cell_tag.push_back(physical);
                    for (int q = 0; q < nn; ++q) m.cell_nodes.push_back(tk.number<int64_t>());
                    continue;
                }
                if (type == 1 || type == 2 || type == 15) {
                    for (int q = 0; q < nn; ++q) (void)tk.number<int64_t>();
                    continue;
                }
                throw ParseError("unsupported mesh topology: SimiLie currently requires the whole mesh to be made "
                                 "of quadrilaterals or hexahedra");
            }
        }
    }
    if (m.node_tag.empty()) throw ParseError("the provided mesh does not contain any node");
    if (!saw_elements || m.cell_tag.empty())
        throw ParseError("unsupported mesh topology: SimiLie currently requires the whole mesh to be made of "
                         "quadrilaterals or hexahedra");
    return m;
}

sil_structured_grid_s* build_grid(const Mesh& m)
{
    const int nd = m.cell_dim;
    const size_t nnodes = m.node_tag.size();
    auto* g = new sil_structured_grid_s();
    g->ndim = nd;
    const std::vector<double>* src[3] = {&m.x, &m.y, &m.z};
    run_threads(3, [&](int a) { g->coords[a] = unique_sorted(*src[a]); });
    if (nd == 2) {
        if (g->coords[2].size() != 1) {
            delete g;
            throw ParseError("the quadrilateral mesh nodes do not lie on a single plane");
        }
        g->z_value = g->coords[2][0];
    }
    const size_t n0 = g->coords[0].size(), n1 = g->coords[1].size(), n2 = nd == 3 ? g->coords[2].size() : 1;
    const size_t total = n0 * n1 * n2;
    // node tag → linear grid index (dense table when the tags are dense)
    int64_t max_tag = 0, min_tag = std::numeric_limits<int64_t>::max();
    for (int64_t t : m.node_tag) {
        max_tag = std::max(max_tag, t);
        min_tag = std::min(min_tag, t);
    }
    const bool dense = min_tag >= 0 && max_tag <= static_cast<int64_t>(4 * nnodes + 16);
    std::vector<int64_t> by_tag_dense(dense ? static_cast<size_t>(max_tag + 1) : 0, -1);
    std::unordered_map<int64_t, int64_t> by_tag;
    if (!dense) by_tag.reserve(nnodes * 2);
    // grid index of every mesh node in parallel (−1: off the axes), then the
    // checks in node order
    std::vector<int64_t> node_lin(nnodes);
    {
        const int T = reader_threads();
        run_threads(T, [&](int t) {
            for (size_t q = nnodes * t / T; q < nnodes * (t + 1) / T; ++q) {
                try {
                    const size_t i = axis_index(g->coords[0], m.x[q]);
                    const size_t j = axis_index(g->coords[1], m.y[q]);
                    const size_t k = nd == 3 ? axis_index(g->coords[2], m.z[q]) : 0;
                    node_lin[q] = static_cast<int64_t>(i + n0 * (j + n1 * k));
                } catch (const std::exception&) {
                    node_lin[q] = -1;
                }
            }
        });
    }
    std::vector<int64_t> owner(total, -1); // mesh node occupying each grid node
    for (size_t q = 0; q < nnodes; ++q) {
        const int64_t lin = node_lin[q];
        if (lin < 0) {
            delete g;
            throw ParseError("coordinate does not belong to the detected structured grid");
        }
        if (owner[lin] >= 0) {
            const int64_t r = owner[lin];
            if (!near(m.x[r], m.x[q]) || !near(m.y[r], m.y[q]) || !near(m.z[r], m.z[q])) {
                delete g;
                throw ParseError("duplicated nodes on the detected rectilinear grid");
            }
        } else {
            owner[lin] = static_cast<int64_t>(q);
        }
        if (dense) {
            if (by_tag_dense[m.node_tag[q]] < 0) by_tag_dense[m.node_tag[q]] = lin;
        } else {
            by_tag.emplace(m.node_tag[q], lin);
        }
    }
    for (int64_t o : owner)
        if (o < 0) {
            delete g;
            throw ParseError("the mesh nodes do not form a full rectilinear grid");
        }
    const size_t c0 = n0 - 1, c1 = n1 - 1, c2 = nd == 3 ? n2 - 1 : 1;
    const size_t ncells = c0 * c1 * c2;
    const char* what = nd == 3 ? "hexahedral" : "quadrilateral";
    const char* one = nd == 3 ? "a hexahedron" : "a quadrilateral";
    if (m.cell_tag.size() != ncells) {
        delete g;
        throw ParseError(std::string("the ") + what + " cells do not cover a full rectilinear grid");
    }
    g->cell_tags.assign(ncells, 0);
    std::vector<char> used(ncells, 0);
    const int per = nd == 3 ? 8 : 4;
    // grid cell of every mesh cell in parallel (−1: an unknown node, −2: not
    // one grid cell), then the checks in cell order
    std::vector<int64_t> cell_lin(ncells);
    {
        const int T = reader_threads();
        run_threads(T, [&](int t) {
            for (size_t c = ncells * t / T; c < ncells * (t + 1) / T; ++c) {
                size_t lo[3] = {SIZE_MAX, SIZE_MAX, SIZE_MAX}, hi[3] = {0, 0, 0};
                int64_t status = 0;
                for (int q = 0; q < per && status == 0; ++q) {
                    const int64_t tag = m.cell_nodes[c * per + q];
                    int64_t lin = -1;
                    if (dense) {
                        if (tag >= 0 && tag <= max_tag) lin = by_tag_dense[tag];
                    } else {
                        auto it = by_tag.find(tag);
                        if (it != by_tag.end()) lin = it->second;
                    }
                    if (lin < 0) {
                        status = -1;
                        break;
                    }
                    const size_t idx[3] = {static_cast<size_t>(lin) % n0, (static_cast<size_t>(lin) / n0) % n1,
                                           static_cast<size_t>(lin) / (n0 * n1)};
                    for (int a = 0; a < 3; ++a) {
                        lo[a] = std::min(lo[a], idx[a]);
                        hi[a] = std::max(hi[a], idx[a]);
                    }
                }
                if (status == 0) {
                    const bool unit = hi[0] == lo[0] + 1 && hi[1] == lo[1] + 1 && (nd == 2 || hi[2] == lo[2] + 1);
                    status = unit ? static_cast<int64_t>(lo[0] + c0 * (lo[1] + c1 * (nd == 3 ? lo[2] : 0))) : -2;
                }
                cell_lin[c] = status;
            }
        });
    }
    for (size_t c = 0; c < ncells; ++c) {
        if (cell_lin[c] == -1) {
            delete g;
            throw ParseError(std::string(one) + " references an unknown node");
        }
        if (cell_lin[c] == -2) {
            delete g;
            throw ParseError(std::string(one) + " does not match a single rectilinear grid cell");
        }
        const size_t lin = static_cast<size_t>(cell_lin[c]);
        if (used[lin]) {
            delete g;
            throw ParseError(std::string("duplicated ") + (nd == 3 ? "hexahedron" : "quadrilateral") +
                             " on the detected rectilinear grid");
        }
        used[lin] = 1;
        g->cell_tags[lin] = m.cell_tag[c];
    }
    if (nd == 2) g->coords[2].clear();
    return g;
}

int read_file(const char* path, std::string& out)
{
    std::ifstream f(path, std::ios::binary);
    SIL_REQUIRE(f.is_open(), SIL_ERR_INVALID, "failed to open file: %s", path);
    f.seekg(0, std::ios::end);
    const std::streamoff size = f.tellg();
    f.seekg(0, std::ios::beg);
    out.resize(static_cast<size_t>(size));
    if (size > 0) f.read(out.data(), size);
    SIL_REQUIRE(f.good() || size == 0, SIL_ERR_INVALID, "failed to read file: %s", path);
    return SIL_OK;
}

int grid_from_text(const char* b, const char* e, sil_structured_grid* out)
{
    try {
        *out = build_grid(parse_msh2(b, e));
        return SIL_OK;
    } catch (const std::bad_alloc&) {
        set_error("msh2: out of host memory");
        return SIL_ERR_INVALID;
    } catch (const std::exception& ex) {
        set_error("%s", ex.what());
        return SIL_ERR_INVALID;
    }
}

// ---------------------------------------------------------------- silpro ---

// Grammar of a .silpro file.  Trivia — whitespace, and comments running
// from '#' or "//" to the end of the line — may stand between any two
// symbols:
//
//   file   := block* END
//   block  := atom '{' member* '}'
//   member := block | atom atom ';'
//   atom   := word | string
//   word   := a first character that is not whitespace, '{', '}', ';', '"'
//             or '#' and does not open "//", then every character up to
//             whitespace, '{', '}', ';' or "//"
//   string := '"' ( '\' any | any but '"' )* '"'   — the backslash keeps the
//             character after it, whatever it is
//
// Only blocks stand at the top level.  Within a block, an atom followed by
// '{' opens a nested block; otherwise the atom is a key and the next atom
// its value.  The first assignment of a key in a block wins, and a block
// whose path was already read is parsed and then dropped whole.
class SilproReader {
  public:
    SilproReader(std::string_view text, sil_silpro_s& out) : s_(text), out_(out) {}

    void file()
    {
        sil_silpro_s::Path path;
        while (true) {
            skip_trivia();
            if (i_ >= s_.size()) return;
            block(path, true);
        }
    }

  private:
    std::string_view s_;
    size_t i_ = 0;
    sil_silpro_s& out_;

    [[noreturn]] void fail(const std::string& what) const
    {
        // 1-based line and column of the reading position
        size_t line = 1, col = 1;
        for (size_t k = 0; k < i_ && k < s_.size(); ++k) {
            if (s_[k] == '\n') {
                ++line;
                col = 1;
            } else {
                ++col;
            }
        }
        throw ParseError(".silpro line " + std::to_string(line) + ", column " + std::to_string(col) + ": " + what);
    }

    bool comment_at(size_t k) const
    {
        return s_[k] == '#' || (s_[k] == '/' && k + 1 < s_.size() && s_[k + 1] == '/');
    }

    static bool blank(char c) { return std::isspace(static_cast<unsigned char>(c)) != 0; }

    void skip_trivia()
    {
        while (i_ < s_.size()) {
            if (blank(s_[i_])) {
                ++i_;
            } else if (comment_at(i_)) {
                const size_t nl = s_.find('\n', i_);
                i_ = nl == std::string_view::npos ? s_.size() : nl;
            } else {
                return;
            }
        }
    }

    // the next symbol after trivia, or 0 at the end of the text
    char upcoming()
    {
        skip_trivia();
        return i_ < s_.size() ? s_[i_] : '\0';
    }

    std::string atom(const std::string& role)
    {
        const char c = upcoming();
        if (i_ >= s_.size()) fail("the file ends where " + role + " should be");
        if (c == '"') {
            std::string text;
            size_t k = i_ + 1;
            while (k < s_.size() && s_[k] != '"') {
                if (s_[k] == '\\' && k + 1 < s_.size()) ++k;
                text.push_back(s_[k++]);
            }
            if (k >= s_.size()) fail("a quoted string is never closed");
            i_ = k + 1;
            return text;
        }
        if (c == '{' || c == '}' || c == ';') fail(std::string("found '") + c + "' where " + role + " should be");
        const size_t begin = i_;
        while (i_ < s_.size() && !blank(s_[i_]) && s_[i_] != '{' && s_[i_] != '}' && s_[i_] != ';' &&
               !(s_[i_] == '/' && i_ + 1 < s_.size() && s_[i_ + 1] == '/'))
            ++i_;
        return std::string(s_.substr(begin, i_ - begin));
    }

    // a block named by the next atom, inside the block at `parent`; `keep`
    // is false inside a block that is dropped
    void block(sil_silpro_s::Path& parent, bool keep)
    {
        parent.push_back(atom("a block name"));
        if (upcoming() != '{') {
            const std::string name = parent.back();
            if (i_ >= s_.size()) fail("the file ends after '" + name + "', where '{' should open its block");
            fail("'" + name + "' must be followed by '{' to open a block");
        }
        ++i_;
        keep = keep && out_.blocks.insert(parent).second;
        while (true) {
            const char c = upcoming();
            if (i_ >= s_.size()) fail("the file ends inside block '" + parent.back() + "'");
            if (c == '}') {
                ++i_;
                break;
            }
            const size_t mark = i_;
            const std::string key = atom("a key or a block name");
            if (upcoming() == '{') {
                i_ = mark; // read again as a block name
                block(parent, keep);
                continue;
            }
            const std::string value = atom("the value of '" + key + "'");
            if (upcoming() != ';') {
                if (i_ >= s_.size()) fail("the file ends before the ';' that ends '" + key + "'");
                fail("the value of '" + key + "' must be followed by ';'");
            }
            ++i_;
            if (keep) {
                parent.push_back(key);
                out_.assignments.try_emplace(parent, value);
                parent.pop_back();
            }
        }
        parent.pop_back();
    }
};

sil_silpro_s::Path dotted(const std::string& path)
{
    sil_silpro_s::Path p;
    size_t b = 0;
    while (true) {
        const size_t e = path.find('.', b);
        p.push_back(path.substr(b, e == std::string::npos ? std::string::npos : e - b));
        if (e == std::string::npos) return p;
        b = e + 1;
    }
}

const std::string* value_at(const sil_silpro_s& f, const sil_silpro_s::Path& key)
{
    auto it = f.assignments.find(key);
    return it == f.assignments.end() ? nullptr : &it->second;
}

// the values assigned directly in the block at `block`, in key order
std::vector<const std::string*> block_values(const sil_silpro_s& f, const sil_silpro_s::Path& block)
{
    std::vector<const std::string*> vals;
    for (auto it = f.assignments.lower_bound(block); it != f.assignments.end(); ++it) {
        const auto& k = it->first;
        if (k.size() < block.size() || !std::equal(block.begin(), block.end(), k.begin())) break;
        if (k.size() == block.size() + 1) vals.push_back(&it->second);
    }
    return vals;
}

// Scalar setting values.  Booleans take the words of the reference's
// parse_bool in any case; numbers must fill the whole value.
bool boolean_setting(const std::string& key, const std::string& v)
{
    static const std::pair<const char*, bool> words[] = {{"1", true},  {"true", true},   {"yes", true},
                                                         {"on", true}, {"0", false},     {"false", false},
                                                         {"no", false}, {"off", false}};
    std::string low;
    for (char c : v)
        if (!std::isspace(static_cast<unsigned char>(c)))
            low.push_back(static_cast<char>(std::tolower(static_cast<unsigned char>(c))));
    for (const auto& [w, b] : words)
        if (low == w) return b;
    throw ParseError(key + ": '" + v + "' is not a boolean (1/0, true/false, yes/no, on/off)");
}

double real_setting(const std::string& key, const std::string& v)
{
    const char* b = v.data();
    const char* e = b + v.size();
    if (b != e && *b == '+') ++b;
    double x = 0.0;
    const auto r = std::from_chars(b, e, x, std::chars_format::general);
    if (b == e || r.ec != std::errc() || r.ptr != e || !std::isfinite(x))
        throw ParseError(key + ": '" + v + "' is not a number");
    return x;
}

uint32_t count_setting(const std::string& key, const std::string& v)
{
    const char* b = v.data();
    const char* e = b + v.size();
    if (b != e && *b == '+') ++b;
    uint32_t x = 0;
    const auto r = std::from_chars(b, e, x, 10);
    if (b == e || r.ec != std::errc() || r.ptr != e)
        throw ParseError(key + ": '" + v + "' is not a non-negative integer that fits 32 bits");
    return x;
}

// accepted spellings of each strong-formulation preconditioner
// (minimize_strong_formulation_residual.hpp:88-117)
int preconditioner_setting(const std::string& v)
{
    static const std::pair<int, std::vector<const char*>> names[] = {
            {SIL_SILPRO_PRECOND_IDENTITY, {"Identity", "identity"}},
            {SIL_SILPRO_PRECOND_JACOBI, {"Jacobi", "jacobi"}},
            {SIL_SILPRO_PRECOND_SPD_ISAI, {"SpdIsai", "spd-isai", "spd_isai"}},
            {SIL_SILPRO_PRECOND_SYMMETRIC_GAUSS_SEIDEL,
             {"SymmetricGaussSeidel", "symmetric-gauss-seidel", "symmetric_gauss_seidel"}},
            {SIL_SILPRO_PRECOND_SSOR, {"Ssor", "SSOR", "ssor"}},
            {SIL_SILPRO_PRECOND_CHEBYSHEV_JACOBI, {"ChebyshevJacobi", "chebyshev-jacobi", "chebyshev_jacobi"}},
            {SIL_SILPRO_PRECOND_IR_JACOBI, {"IrJacobi", "ir-jacobi", "ir_jacobi"}},
            {SIL_SILPRO_PRECOND_GENERAL_ISAI, {"GeneralIsai", "general-isai", "general_isai", "isai"}},
    };
    for (const auto& [id, spellings] : names)
        for (const char* n : spellings)
            if (v == n) return id;
    throw ParseError("Preconditioner: '" + v + "' names no strong-formulation preconditioner");
}

int enum_setting(const std::string& key, const std::string& v, std::initializer_list<const char*> names)
{
    int id = 0;
    for (const char* n : names) {
        if (v == n) return id;
        ++id;
    }
    throw ParseError(key + ": '" + v + "' is not supported");
}

int copy_out(const std::string& v, char* buf, int64_t size)
{
    SIL_REQUIRE(buf && size > 0, SIL_ERR_INVALID, "silpro: no output buffer");
    SIL_REQUIRE(static_cast<int64_t>(v.size()) < size, SIL_ERR_INVALID, "silpro: value of %zu bytes does not fit",
                v.size());
    std::memcpy(buf, v.c_str(), v.size() + 1);
    return SIL_OK;
}

}  // namespace
}  // namespace sil

using namespace sil;

extern "C" {

int sil_msh2_read(const char* path, sil_structured_grid* out)
{
    SIL_REQUIRE(path && out, SIL_ERR_INVALID, "msh2_read: null argument");
    std::string text;
    int rc = read_file(path, text);
    if (rc) return rc;
    return grid_from_text(text.data(), text.data() + text.size(), out);
}

int sil_msh2_parse(const char* text, int64_t length, sil_structured_grid* out)
{
    SIL_REQUIRE(text && out && length >= 0, SIL_ERR_INVALID, "msh2_parse: null argument");
    return grid_from_text(text, text + length, out);
}

int sil_structured_grid_shape(sil_structured_grid g, int32_t* ndim, int64_t* nodes)
{
    SIL_REQUIRE(g && ndim && nodes, SIL_ERR_INVALID, "structured_grid_shape: null argument");
    *ndim = g->ndim;
    for (int a = 0; a < 3; ++a) nodes[a] = a < g->ndim ? static_cast<int64_t>(g->coords[a].size()) : 1;
    return SIL_OK;
}

int sil_structured_grid_coords(sil_structured_grid g, int32_t axis, double* out)
{
    SIL_REQUIRE(g && out && axis >= 0 && axis < g->ndim, SIL_ERR_INVALID, "structured_grid_coords: bad argument");
    std::copy(g->coords[axis].begin(), g->coords[axis].end(), out);
    return SIL_OK;
}

double sil_structured_grid_plane(sil_structured_grid g) { return g ? g->z_value : 0.0; }

int sil_structured_grid_cell_tags(sil_structured_grid g, int32_t* out)
{
    SIL_REQUIRE(g && out, SIL_ERR_INVALID, "structured_grid_cell_tags: null argument");
    std::copy(g->cell_tags.begin(), g->cell_tags.end(), out);
    return SIL_OK;
}

int sil_structured_grid_destroy(sil_structured_grid g)
{
    delete g;
    return SIL_OK;
}

int sil_silpro_parse(const char* text, int64_t length, sil_silpro* out)
{
    SIL_REQUIRE(text && out && length >= 0, SIL_ERR_INVALID, "silpro_parse: null argument");
    try {
        auto* h = new sil_silpro_s();
        try {
            SilproReader(std::string_view(text, static_cast<size_t>(length)), *h).file();
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return SIL_OK;
    } catch (const std::exception& ex) {
        set_error("%s", ex.what());
        return SIL_ERR_INVALID;
    }
}

int sil_silpro_read(const char* path, sil_silpro* out)
{
    SIL_REQUIRE(path && out, SIL_ERR_INVALID, "silpro_read: null argument");
    std::string text;
    int rc = read_file(path, text);
    if (rc) return rc;
    return sil_silpro_parse(text.data(), static_cast<int64_t>(text.size()), out);
}

int sil_silpro_value(sil_silpro h, const char* path, char* buf, int64_t size)
{
    SIL_REQUIRE(h && path, SIL_ERR_INVALID, "silpro_value: null argument");
    const std::string* v = value_at(*h, dotted(path));
    SIL_REQUIRE(v != nullptr, SIL_ERR_INVALID, "silpro: no value '%s'", path);
    return copy_out(*v, buf, size);
}

int sil_silpro_section_value(sil_silpro h, const char* section, int32_t index, char* buf, int64_t size)
{
    SIL_REQUIRE(h && section && index >= 0, SIL_ERR_INVALID, "silpro_section_value: bad argument");
    const sil_silpro_s::Path block = dotted(section);
    SIL_REQUIRE(h->blocks.count(block) != 0, SIL_ERR_INVALID, "silpro: no section '%s'", section);
    const auto vals = block_values(*h, block);
    SIL_REQUIRE(index < static_cast<int32_t>(vals.size()), SIL_ERR_INVALID, "silpro: '%s' has %zu values", section,
                vals.size());
    return copy_out(*vals[static_cast<size_t>(index)], buf, size);
}

int sil_silpro_problem_settings(sil_silpro h, sil_silpro_settings* out)
{
    SIL_REQUIRE(h && out, SIL_ERR_INVALID, "silpro_problem_settings: null argument");
    try {
        for (const char* b : {"Problem", "Solver"})
            if (!h->blocks.count({b})) throw ParseError(std::string("the file has no '") + b + "' block");
        // the value of Block.key, or the default when the block does not set it
        auto get = [&](const char* block, const char* key, const char* dflt) -> std::string {
            const std::string* v = value_at(*h, {block, key});
            return v ? *v : std::string(dflt);
        };
        sil_silpro_settings st{};
        st.physics = enum_setting("Physics", get("Problem", "Physics", "LinearMagnetostatics"),
                                  {"LinearMagnetostatics", "NonLinearMagnetostatics", "ScalarFieldWithPowerCoupling"});
        static_assert(SIL_PHYSICS_LINEAR_MAGNETOSTATICS == 0 && SIL_PHYSICS_NONLINEAR_MAGNETOSTATICS == 1 &&
                      SIL_PHYSICS_SCALAR_FIELD_WITH_POWER_COUPLING == 2);
        enum_setting("Problem.Solver", get("Problem", "Solver", "MinimizeStrongFormulationResidual"),
                     {"MinimizeStrongFormulationResidual"});
        // MinimizeStrongFormulationResidualProblem defaults (onelab_interface.hpp:96-111)
        st.max_iterations = count_setting("MaxIterations", get("Solver", "MaxIterations", "10000"));
        st.relative_tolerance = real_setting("RelativeTolerance", get("Solver", "RelativeTolerance", "1e-10"));
        st.jacobi_max_block_size =
                static_cast<int32_t>(count_setting("JacobiMaxBlockSize", get("Solver", "JacobiMaxBlockSize", "1")));
        st.use_matrix_free = boolean_setting("UseMatrixFree", get("Solver", "UseMatrixFree", "1")) ? 1 : 0;
        st.criterion = enum_setting("Criterion", get("Solver", "Criterion", "MomentsTemporalDerivative"),
                                    {"PotentialTemporalDerivative", "MomentsTemporalDerivative",
                                     "PotentialAndMomentsTemporalDerivative"});
        st.preconditioner = preconditioner_setting(get("Solver", "Preconditioner", "Jacobi"));
        st.vector_potential_gauge_penalty =
                real_setting("VectorPotentialGaugePenalty", get("Solver", "VectorPotentialGaugePenalty", "20"));
        st.sor_relaxation_factor = real_setting("SorRelaxationFactor", get("Solver", "SorRelaxationFactor", "1.2"));
        st.chebyshev_iterations =
                static_cast<int32_t>(count_setting("ChebyshevIterations", get("Solver", "ChebyshevIterations", "8")));
        st.chebyshev_lower_bound = real_setting("ChebyshevLowerBound", get("Solver", "ChebyshevLowerBound", "0.05"));
        st.chebyshev_upper_bound = real_setting("ChebyshevUpperBound", get("Solver", "ChebyshevUpperBound", "1.5"));
        st.ir_iterations = static_cast<int32_t>(count_setting("IrIterations", get("Solver", "IrIterations", "8")));
        st.ir_relaxation_factor = real_setting("IrRelaxationFactor", get("Solver", "IrRelaxationFactor", "1"));
        *out = st;
        return SIL_OK;
    } catch (const std::exception& ex) {
        set_error("%s", ex.what());
        return SIL_ERR_INVALID;
    }
}

int sil_silpro_solver_settings(const sil_silpro_settings* in, sil_solver_settings* out)
{
    SIL_REQUIRE(in && out, SIL_ERR_INVALID, "silpro_solver_settings: null argument");
    SIL_REQUIRE(in->max_iterations <= INT32_MAX, SIL_ERR_INVALID, "MaxIterations %lld exceeds the int32 range",
                (long long)in->max_iterations);
    SIL_REQUIRE(in->jacobi_max_block_size == 1, SIL_ERR_UNSUPPORTED,
                "JacobiMaxBlockSize %d: only scalar Jacobi (block size 1) is implemented", in->jacobi_max_block_size);
    int pc;
    switch (in->preconditioner) {
    case SIL_SILPRO_PRECOND_IDENTITY: pc = SIL_PRECOND_IDENTITY; break;
    case SIL_SILPRO_PRECOND_JACOBI: pc = SIL_PRECOND_JACOBI; break;
    case SIL_SILPRO_PRECOND_SPD_ISAI: pc = SIL_PRECOND_SPD_ISAI; break;
    default:
        set_error("preconditioner %d (GaussSeidel / SSOR / Chebyshev / IR / GeneralIsai) is not implemented",
                  in->preconditioner);
        return SIL_ERR_UNSUPPORTED;
    }
    sil_default_solver_settings(out);
    out->max_iterations = static_cast<int32_t>(in->max_iterations);
    out->relative_tolerance = in->relative_tolerance;
    out->preconditioner = pc;
    return SIL_OK;
}

int sil_silpro_destroy(sil_silpro h)
{
    delete h;
    return SIL_OK;
}

int sil_magnetostatics_cell_inputs(const int32_t* cell_tags, int64_t ncells, const sil_cell_materials* mat,
                                   double* cell_mu, double* cell_current, double* nonlinear_flags)
{
    SIL_REQUIRE(cell_tags && mat && cell_mu && cell_current && nonlinear_flags, SIL_ERR_INVALID,
                "magnetostatics_cell_inputs: null argument");
    auto has = [](const int32_t* tags, int32_t n, int32_t t) {
        for (int32_t i = 0; i < n; ++i)
            if (tags[i] == t) return true;
        return false;
    };
    for (int64_t c = 0; c < ncells; ++c) {
        const int32_t t = cell_tags[c];
        double mu = mat->mu0, j[3] = {0.0, 0.0, 0.0}, nl = 0.0;
        if (has(mat->magnetic_material_tags, mat->n_magnetic_material_tags, t)) {
            mu = mat->core_mu;
            nl = mat->use_nonlinear_magnetic_material ? 1.0 : 0.0;
        } else if (has(mat->positive_x_tags, mat->n_positive_x_tags, t)) {
            j[0] = mat->current_density_x;
        } else if (has(mat->negative_x_tags, mat->n_negative_x_tags, t)) {
            j[0] = -mat->current_density_x;
        } else if (has(mat->positive_z_tags, mat->n_positive_z_tags, t)) {
            j[2] = mat->current_density_z;
        } else if (has(mat->negative_z_tags, mat->n_negative_z_tags, t)) {
            j[2] = -mat->current_density_z;
        }
        cell_mu[c] = mu;
        for (int a = 0; a < 3; ++a) cell_current[3 * c + a] = j[a];
        nonlinear_flags[c] = nl;
    }
    return SIL_OK;
}

}  // extern "C"
