// General per-node kernels of the sil::exterior operators: any dimension
// N ≤ 4, any degree k, strided index-last / index-first views, spectator
// dimensions (coboundary), rectilinear or nodal positions and identity /
// diagonal / symmetric metrics.  One thread per output node; every
// intermediate (dω, ⋆, dᵀ, δω) is re-evaluated where it is needed, so no
// workspace is ever allocated.  The fused kernels in dec_fused.hip serve the
// hot shapes; these serve everything else.
#include "dec_kernels.hpp"
#include "geometry.hpp"

namespace sil {

namespace {

struct Coord {
    int64_t i[4];
};

__device__ inline void decode(const DevView& v, int64_t t, Coord& node, Coord& batch)
{
    for (int a = v.nbatch - 1; a >= 0; --a) {
        batch.i[a] = t % v.bext[a];
        t /= v.bext[a];
    }
    for (int a = v.ndim - 1; a >= 0; --a) {
        node.i[a] = t % v.ext[a];
        t /= v.ext[a];
    }
}

__device__ inline int64_t batch_offset(const DevView& v, const Coord& b)
{
    int64_t off = 0;
    for (int a = 0; a < v.nbatch; ++a) off += b.i[a] * v.bstr[a];
    return off;
}

__device__ inline int64_t node_offset(const DevView& v, const int64_t* idx)
{
    int64_t off = 0;
    for (int a = 0; a < v.ndim; ++a) off += idx[a] * v.str[a];
    return off;
}

// ---------------------------------------------------------- coboundary ---

// (dω)_A(n) = Σ_m (−1)^m [ω_{A∖a_m}(clamp(n+e_{a_m})) − ω_{A∖a_m}(clamp(n))]
// transposed: Σ_m (−1)^m [η_{A∖a_m}(n) − η_{A∖a_m}(n−e_{a_m})], zero outside
// (coboundary.hpp:353-427, 517-605 through boundary.hpp:43-65, 174-189).
__global__ void k_coboundary(int k, int transposed, const double* __restrict__ in, DevView vin,
                             double* __restrict__ out, DevView vout, int64_t total)
{
    const int n = vout.ndim;
    const int nup = c_combos.count[n][k + 1];
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        Coord node, batch;
        decode(vout, t, node, batch);
        const double* inb = in + batch_offset(vin, batch);
        double* outb = out + batch_offset(vout, batch);
        int64_t base[4];
        for (int a = 0; a < n; ++a) base[a] = transposed ? node.i[a] : min(node.i[a], vin.ext[a] - 1);
        for (int c = 0; c < nup; ++c) {
            unsigned A = c_combos.mask[n][k + 1][c];
            double acc = 0.0;
            int m = 0;
            for (int a = 0; a < n; ++a) {
                if (!(A & (1u << a))) continue;
                int li = c_combos.index[n][A & ~(1u << a)];
                double sgn = (m & 1) ? -1.0 : 1.0;
                ++m;
                int64_t s[4];
                for (int q = 0; q < n; ++q) s[q] = base[q];
                if (!transposed) {
                    s[a] = min(s[a] + 1, vin.ext[a] - 1);
                    double fwd = inb[node_offset(vin, s) + li * vin.cstr];
                    double cur = inb[node_offset(vin, base) + li * vin.cstr];
                    acc += sgn * (fwd - cur);
                } else {
                    bool in_here = true;
                    for (int q = 0; q < n; ++q) in_here = in_here && base[q] < vin.ext[q];
                    double cur = in_here ? inb[node_offset(vin, base) + li * vin.cstr] : 0.0;
                    s[a] -= 1;
                    bool in_back = s[a] >= 0;
                    for (int q = 0; q < n; ++q) in_back = in_back && s[q] < vin.ext[q];
                    double back = in_back ? inb[node_offset(vin, s) + li * vin.cstr] : 0.0;
                    acc += sgn * (cur - back);
                }
            }
            outb[node_offset(vout, node.i) + c * vout.cstr] = acc;
        }
    }
}

// ----------------------------------------------------------- Hodge star ---

__global__ void k_fill_hodge(DevGeom g, int k, double* __restrict__ out, int64_t nodes)
{
    const int n = g.ndim;
    int64_t nfull = 1;
    for (int i = 0; i < k; ++i) nfull *= n;
    const int nq = c_combos.count[n][n - k];
    double fact = 1.0;
    for (int i = 2; i <= k; ++i) fact *= i;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nodes;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t idx[4];
        int64_t r = t;
        for (int a = n - 1; a >= 0; --a) {
            idx[a] = r % g.ext[a];
            r /= g.ext[a];
        }
        NodeGeo ng;
        load_node_geo(g, idx, ng);
        for (int64_t f = 0; f < nfull; ++f) {
            int tup[4];
            int64_t rr = f;
            for (int i = k - 1; i >= 0; --i) {
                tup[i] = (int)(rr % n);
                rr /= n;
            }
            unsigned src = 0;
            bool unique = true;
            int odd = 0;
            for (int i = 0; i < k; ++i) {
                if (src & (1u << tup[i])) unique = false;
                src |= 1u << tup[i];
                for (int j = i + 1; j < k; ++j)
                    if (tup[i] > tup[j]) odd ^= 1;
            }
            double coef = unique ? node_hodge(ng, src) : 0.0;
            unsigned tgt_of_src = ((1u << n) - 1u) & ~src;
            for (int q = 0; q < nq; ++q) {
                unsigned tgt = c_combos.mask[n][n - k][q];
                double v = (unique && tgt == tgt_of_src) ? (odd ? -1.0 : 1.0) / fact * coef : 0.0;
                out[(t * nfull + f) * nq + q] = v;
            }
        }
    }
}

__global__ void k_hodge_apply(DevGeom g, int k, const double* __restrict__ in, DevView vin,
                              double* __restrict__ out, DevView vout, int64_t nodes)
{
    const int n = g.ndim;
    const int nq = c_combos.count[n][n - k];
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nodes;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t idx[4];
        int64_t r = t;
        for (int a = n - 1; a >= 0; --a) {
            idx[a] = r % g.ext[a];
            r /= g.ext[a];
        }
        NodeGeo ng;
        load_node_geo(g, idx, ng);
        for (int q = 0; q < nq; ++q) {
            unsigned tgt = c_combos.mask[n][n - k][q];
            unsigned src = ((1u << n) - 1u) & ~tgt;
            int li = c_combos.index[n][src];
            double w = in[node_offset(vin, idx) + li * vin.cstr];
            out[node_offset(vout, idx) + q * vout.cstr] = node_hodge(ng, src) * w;
        }
    }
}

// ------------------------------------------------- on-the-fly compositions ---

// ω component with axes set `mask` at node idx (zero outside when zero_out)
__device__ inline double sample(const double* in, const DevView& v, const int64_t* idx, unsigned mask, int n,
                                bool zero_out)
{
    int64_t s[4];
    for (int a = 0; a < n; ++a) {
        s[a] = idx[a];
        if (s[a] < 0 || s[a] >= v.ext[a]) {
            if (zero_out) return 0.0;
            s[a] = s[a] < 0 ? 0 : v.ext[a] - 1;
        }
    }
    int li = c_combos.index[n][mask];
    return in[node_offset(v, s) + li * v.cstr];
}

// (dω)_A at node p, clamped into the input domain
__device__ inline double d_at(const double* in, const DevView& v, const int64_t* p, unsigned A, int n)
{
    int64_t base[4];
    for (int a = 0; a < n; ++a) base[a] = min(max(p[a], (int64_t)0), v.ext[a] - 1);
    double acc = 0.0;
    int m = 0;
    for (int a = 0; a < n; ++a) {
        if (!(A & (1u << a))) continue;
        double sgn = (m & 1) ? -1.0 : 1.0;
        ++m;
        int64_t s[4];
        for (int q = 0; q < n; ++q) s[q] = base[q];
        s[a] = min(s[a] + 1, v.ext[a] - 1);
        acc += sgn * (sample(in, v, s, A & ~(1u << a), n, false) - sample(in, v, base, A & ~(1u << a), n, false));
    }
    return acc;
}

// source form of a codifferential: either ω itself or dω (on-the-fly)
struct FormSource {
    const double* in;
    DevView v;
    bool is_d; // value = (dω)_mask instead of ω_mask
};

__device__ inline double form_value(const FormSource& f, const int64_t* p, unsigned mask, int n)
{
    if (f.is_d) return d_at(f.in, f.v, p, mask, n);
    return sample(f.in, f.v, p, mask, n, false);
}

// (δβ)_{I'}(node) for a kq-form β given by `src` on the full domain:
// s·⋆_dual( dᵀ( ⋆β ) ), codifferential.hpp:232-310
__device__ inline double codiff_at(const DevGeom& g, const FormSource& src, int kq, const int64_t* node,
                                   unsigned Iprime)
{
    const int n = g.ndim;
    const unsigned all = (1u << n) - 1u;
    const unsigned Idd = all & ~Iprime; // (N-kq+1)-set
    double G = 0.0;
    int m = 0;
    for (int a = 0; a < n; ++a) {
        if (!(Idd & (1u << a))) continue;
        double sgn = (m & 1) ? -1.0 : 1.0;
        ++m;
        unsigned J = Idd & ~(1u << a);   // dual (N-kq)-set
        unsigned srcset = all & ~J;      // kq-set = I' ∪ {a}
        // η_J at node and node - e_a (zero outside the node domain)
        double eta_here, eta_back;
        {
            NodeGeo ng;
            load_node_geo(g, node, ng);
            eta_here = node_hodge(ng, srcset) * form_value(src, node, srcset, n);
        }
        int64_t b[4];
        for (int q = 0; q < n; ++q) b[q] = node[q];
        b[a] -= 1;
        if (b[a] >= 0) {
            NodeGeo ng;
            load_node_geo(g, b, ng);
            eta_back = node_hodge(ng, srcset) * form_value(src, b, srcset, n);
        } else {
            eta_back = 0.0;
        }
        G += sgn * (eta_here - eta_back);
    }
    NodeGeo ng;
    load_node_geo(g, node, ng);
    double s = ((n * (kq + 1) + 1) % 2 == 1) ? -1.0 : 1.0;
    return s * node_hodge(ng, Idd) * G;
}

__global__ void k_codifferential(DevGeom g, int k, const double* __restrict__ in, DevView vin,
                                 double* __restrict__ out, DevView vout, int64_t total)
{
    const int n = g.ndim;
    const int no = c_combos.count[n][k - 1];
    FormSource src{in, vin, false};
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t idx[4];
        int64_t r = t;
        for (int a = n - 1; a >= 0; --a) {
            idx[a] = r % vout.ext[a];
            r /= vout.ext[a];
        }
        for (int q = 0; q < no; ++q) {
            unsigned Ip = c_combos.mask[n][k - 1][q];
            out[node_offset(vout, idx) + q * vout.cstr] = codiff_at(g, src, k, idx, Ip);
        }
    }
}

// Hodge Laplacian at one output node (laplacian.hpp:27-155, 599-647, 791-811)
__global__ void k_laplacian(DevGeom g, int k, const double* __restrict__ in, DevView vin,
                            double* __restrict__ out, DevView vout, int64_t total)
{
    const int n = g.ndim;
    const int nk = c_combos.count[n][k];
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t idx[4];
        int64_t r = t;
        for (int a = n - 1; a >= 0; --a) {
            idx[a] = r % vout.ext[a];
            r /= vout.ext[a];
        }
        for (int q = 0; q < nk; ++q) {
            unsigned I = c_combos.mask[n][k][q];
            double acc = 0.0;
            if (k < n) {
                FormSource dsrc{in, vin, true};
                acc += codiff_at(g, dsrc, k + 1, idx, I);
            }
            if (k > 0) {
                // (d δω)_I(n): δω on the full domain, d clamped into it
                FormSource wsrc{in, vin, false};
                int64_t base[4];
                for (int a = 0; a < n; ++a) base[a] = min(idx[a], g.ext[a] - 1);
                int m = 0;
                for (int a = 0; a < n; ++a) {
                    if (!(I & (1u << a))) continue;
                    double sgn = (m & 1) ? -1.0 : 1.0;
                    ++m;
                    unsigned Ip = I & ~(1u << a);
                    int64_t f[4];
                    for (int q2 = 0; q2 < n; ++q2) f[q2] = base[q2];
                    f[a] = min(f[a] + 1, g.ext[a] - 1);
                    acc += sgn * (codiff_at(g, wsrc, k, f, Ip) - codiff_at(g, wsrc, k, base, Ip));
                }
            }
            out[node_offset(vout, idx) + q * vout.cstr] = acc;
        }
    }
}

inline unsigned grid_for(int64_t total, int block)
{
    int64_t b = (total + block - 1) / block;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return static_cast<unsigned>(b);
}

}  // namespace

int launch_coboundary(int k, bool transposed, const double* in, const sil_view& vin, double* out,
                      const sil_view& vout, hipStream_t stream)
{
    int64_t total = view_nodes(vout) * view_batches(vout);
    if (total == 0) return SIL_OK;
    const int block = 256;
    hipLaunchKernelGGL(k_coboundary, dim3(grid_for(total, block)), dim3(block), 0, stream, k, transposed ? 1 : 0, in,
                       to_dev(vin), out, to_dev(vout), total);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int launch_fill_hodge(const DevGeom& g, int k, double* out, hipStream_t stream)
{
    int64_t nodes = 1;
    for (int a = 0; a < g.ndim; ++a) nodes *= g.ext[a];
    const int block = 128;
    hipLaunchKernelGGL(k_fill_hodge, dim3(grid_for(nodes, block)), dim3(block), 0, stream, g, k, out, nodes);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int launch_hodge_apply(const DevGeom& g, int k, const double* in, const sil_view& vin, double* out,
                       const sil_view& vout, hipStream_t stream)
{
    int64_t nodes = 1;
    for (int a = 0; a < g.ndim; ++a) nodes *= g.ext[a];
    const int block = 128;
    hipLaunchKernelGGL(k_hodge_apply, dim3(grid_for(nodes, block)), dim3(block), 0, stream, g, k, in, to_dev(vin),
                       out, to_dev(vout), nodes);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int launch_codifferential(const DevGeom& g, int k, const double* in, const sil_view& vin, double* out,
                          const sil_view& vout, hipStream_t stream)
{
    int64_t total = view_nodes(vout);
    if (total == 0) return SIL_OK;
    const int block = 128;
    hipLaunchKernelGGL(k_codifferential, dim3(grid_for(total, block)), dim3(block), 0, stream, g, k, in, to_dev(vin),
                       out, to_dev(vout), total);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

int launch_laplacian_generic(const DevGeom& g, int k, const double* in, const sil_view& vin, double* out,
                             const sil_view& vout, hipStream_t stream)
{
    int64_t total = view_nodes(vout);
    if (total == 0) return SIL_OK;
    const int block = 128;
    hipLaunchKernelGGL(k_laplacian, dim3(grid_for(total, block)), dim3(block), 0, stream, g, k, in, to_dev(vin), out,
                       to_dev(vout), total);
    SIL_LAUNCH_CHECK();
    return SIL_OK;
}

}  // namespace sil
