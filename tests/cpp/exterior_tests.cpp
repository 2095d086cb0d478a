// C++ parity tests of the sil::exterior interface (include/sil/exterior.hpp),
// written like the reference's GTest suite (tests/exterior/*.cpp): the same
// grids, inputs and expected values, run on the GPU through the C ABI.
// Built by `python -m similie_amd.build` into tests/cpp/exterior_tests and run
// by tests/test_gpu_cpp.py.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "sil/exterior.hpp"

namespace {

struct Registry {
    std::vector<std::pair<std::string, std::function<void()>>> tests;
    int failures = 0;
    int checks = 0;
};
Registry& reg()
{
    static Registry r;
    return r;
}
struct Add {
    Add(const char* n, std::function<void()> f) { reg().tests.emplace_back(n, std::move(f)); }
};
#define TEST(suite, name)                                         \
    static void suite##_##name();                                 \
    static Add add_##suite##_##name(#suite "." #name, suite##_##name); \
    static void suite##_##name()

#define EXPECT_EQ(a, b)                                                                            \
    do {                                                                                           \
        ++reg().checks;                                                                            \
        double va_ = (a), vb_ = (b);                                                               \
        if (!(va_ == vb_)) {                                                                       \
            std::printf("  FAIL %s:%d: %s == %s (%.17g vs %.17g)\n", __FILE__, __LINE__, #a, #b, va_, vb_); \
            ++reg().failures;                                                                      \
        }                                                                                          \
    } while (0)

#define EXPECT_NEAR(a, b, tol)                                                                     \
    do {                                                                                           \
        ++reg().checks;                                                                            \
        double va_ = (a), vb_ = (b);                                                               \
        if (!(std::fabs(va_ - vb_) <= (tol))) {                                                    \
            std::printf("  FAIL %s:%d: |%s - %s| <= %g (%.17g vs %.17g)\n", __FILE__, __LINE__, #a, #b, \
                        (double)(tol), va_, vb_);                                                  \
            ++reg().failures;                                                                      \
        }                                                                                          \
    } while (0)

struct Dev {
    double* p = nullptr;
    size_t n = 0;
    explicit Dev(const std::vector<double>& h) : n(h.size())
    {
        if (hipMalloc(&p, sizeof(double) * n) != hipSuccess) throw std::runtime_error("hipMalloc");
        if (hipMemcpy(p, h.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
            throw std::runtime_error("hipMemcpy");
    }
    explicit Dev(size_t count) : n(count)
    {
        if (hipMalloc(&p, sizeof(double) * n) != hipSuccess) throw std::runtime_error("hipMalloc");
    }
    ~Dev() { (void)hipFree(p); }
    std::vector<double> host() const
    {
        std::vector<double> h(n);
        (void)hipDeviceSynchronize();
        if (hipMemcpy(h.data(), p, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
            throw std::runtime_error("hipMemcpy");
        return h;
    }
};

// UniformPointSampling::init(a, b, n): n points with step (b − a)/(n − 1)
std::vector<double> uniform(double a, double b, int n)
{
    std::vector<double> c(n);
    for (int i = 0; i < n; ++i) c[i] = a + (b - a) * i / (n - 1);
    return c;
}

using namespace sil;

// tests/exterior/derivative.cpp:138-181 (index-last and index-first)
TEST(ExteriorDerivative, 1DGradient)
{
    for (bool coalescent : {false, true}) {
        std::vector<double> in = {1.0, 2.0, 1.0};
        Dev din(in), dout(3);
        auto vi = coalescent ? index_first<1>(din.p, {3}) : index_last<1>(din.p, {3}, 1);
        auto vo = coalescent ? index_first<1>(dout.p, {3}) : index_last<1>(dout.p, {3}, 1);
        exterior::deriv<1>(nullptr, vo, vi, 0);
        auto o = dout.host();
        EXPECT_EQ(o[0], 1.);
        EXPECT_EQ(o[1], -1.);
        EXPECT_EQ(o[2], 0.);
    }
}

// tests/exterior/derivative.cpp:381-437
TEST(ExteriorDerivative, 2DRotational)
{
    const int N = 3;
    std::vector<double> in(N * N * 2);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int c = 0; c < 2; ++c) in[(i * N + j) * 2 + c] = c + 1.0 + ((i == 1 && j == 1) ? 1.0 : 0.0);
    Dev din(in), dout(4);
    exterior::deriv<2>(nullptr, index_last<2>(dout.p, {2, 2}, 1), index_last<2>(din.p, {N, N}, 2), 1);
    auto o = dout.host();
    EXPECT_EQ(o[0 * 2 + 0], 0.);
    EXPECT_EQ(o[1 * 2 + 0], -1.);
    EXPECT_EQ(o[0 * 2 + 1], 1.);
    EXPECT_EQ(o[1 * 2 + 1], 0.);
}

// tests/exterior/derivative.cpp:654-717
TEST(ExteriorDerivative, 3DDivergency)
{
    const int N = 3;
    std::vector<double> in(N * N * N * 3);
    for (int n = 0; n < N * N * N; ++n)
        for (int c = 0; c < 3; ++c) in[n * 3 + c] = c + 1.0 + (n == (1 * N + 1) * N + 1 ? 1.0 : 0.0);
    Dev din(in), dout(8);
    exterior::deriv<3>(nullptr, index_last<3>(dout.p, {2, 2, 2}, 1), index_last<3>(din.p, {N, N, N}, 3), 2);
    auto o = dout.host();
    EXPECT_EQ(o[(0 * 2 + 1) * 2 + 1], 1.);
    EXPECT_EQ(o[(1 * 2 + 0) * 2 + 1], -1.);
    EXPECT_EQ(o[(1 * 2 + 1) * 2 + 0], 1.);
    EXPECT_EQ(o[(1 * 2 + 1) * 2 + 1], -1.);
}

// tests/exterior/hodge_star.cpp:65-181
TEST(DiscreteHodgeStar, Metric3D)
{
    std::vector<double> metric;
    for (int n = 0; n < 8; ++n)
        for (double v : {4., 1., 2., 5., 3., 6.}) metric.push_back(v);
    Dev dm(metric);
    auto axis = uniform(0., 2., 2);
    exterior::Grid<3> grid({axis, axis, axis}, SIL_METRIC_SYMMETRIC, dm.p);
    Dev hodge(8 * 9 * 3), hodge2(8 * 3 * 3);
    exterior::fill_discrete_hodge_star<3>(nullptr, grid, 2, hodge.p);
    exterior::fill_discrete_hodge_star<3>(nullptr, grid, 1, hodge2.p);
    auto h = hodge.host(), h2 = hodge2.host();
    for (int n = 0; n < 8; ++n) {
        // form = 3 dx∧dy: natural (X,Y) = 3, (Y,X) = −3
        double dual_z = h[(n * 9 + 1) * 3 + 2] * 3. + h[(n * 9 + 3) * 3 + 2] * -3.;
        EXPECT_NEAR(dual_z, 1.5 * std::sqrt(6. / 19.), 4 * 2.2e-16 * 1.5 * std::sqrt(6. / 19.));
        double back_xy = h2[(n * 3 + 2) * 3 + 0] * dual_z;
        EXPECT_NEAR(back_xy, 3., 4 * 2.2e-16 * 3.);
    }
}

// tests/exterior/codifferential.cpp:40-117
TEST(Codifferential, NonStaged2D1Form)
{
    const int n = 31;
    auto c = uniform(-5., 5., n);
    const double dx = 10. / 31.;
    std::vector<double> form(n * n * 2), metric(n * n * 3);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            form[(i * n + j) * 2 + 0] = c[i] * dx;
            form[(i * n + j) * 2 + 1] = c[j] * dx;
            metric[(i * n + j) * 3 + 0] = 1.;
            metric[(i * n + j) * 3 + 1] = 0.;
            metric[(i * n + j) * 3 + 2] = 1.;
        }
    Dev df(form), dm(metric), dout(n * n);
    exterior::Grid<2> grid({c, c}, SIL_METRIC_SYMMETRIC, dm.p);
    exterior::codifferential<2>(nullptr, grid, index_last<2>(dout.p, {n, n}, 1), index_last<2>(df.p, {n, n}, 2), 1);
    auto o = dout.host();
    const double expected = -30. / 31. - 30. / 31.;
    for (int i = 1; i < n - 1; ++i)
        for (int j = 1; j < n - 1; ++j) EXPECT_NEAR(o[i * n + j], expected, 1e-12);
}

// codifferential.hpp:321-508: the staged codifferential gives the closed
// form of the reference's codifferential test and the one-shot result bit
// for bit, run twice on one handle
TEST(Codifferential, Staged2D1Form)
{
    const int n = 31;
    auto c = uniform(-5., 5., n);
    const double dx = 10. / 31.;
    std::vector<double> form(n * n * 2), metric(n * n * 3);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            form[(i * n + j) * 2 + 0] = c[i] * dx;
            form[(i * n + j) * 2 + 1] = c[j] * dx;
            metric[(i * n + j) * 3 + 0] = 1.;
            metric[(i * n + j) * 3 + 1] = 0.;
            metric[(i * n + j) * 3 + 2] = 1.;
        }
    Dev df(form), dm(metric), d1(n * n), d2(n * n);
    exterior::Grid<2> grid({c, c}, SIL_METRIC_SYMMETRIC, dm.p);
    auto in = index_last<2>(df.p, {n, n}, 2);
    auto o1 = index_last<2>(d1.p, {n, n}, 1), o2 = index_last<2>(d2.p, {n, n}, 1);
    exterior::codifferential<2>(nullptr, grid, o1, in, 1);
    auto staged = exterior::make_staged_codifferential<2>(grid, o2, in, 1);
    for (int rep = 0; rep < 2; ++rep) {
        staged.run(nullptr, o2, in);
        auto a = d1.host(), b = d2.host();
        for (size_t q = 0; q < a.size(); ++q) EXPECT_EQ(a[q], b[q]);
    }
    auto o = d2.host();
    const double expected = -30. / 31. - 30. / 31.;
    for (int i = 1; i < n - 1; ++i)
        for (int j = 1; j < n - 1; ++j) EXPECT_NEAR(o[i * n + j], expected, 1e-12);
}

// tests/exterior/laplacian.cpp:487-556 (identity metric, output on the domain minus the last nodes)
TEST(Laplacian, 2D0Form)
{
    const int n = 31;
    auto c = uniform(-5., 5., n);
    const double R = 2., alpha = 0.25;
    std::vector<double> pot(n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double r = std::sqrt(c[i] * c[i] + c[j] * c[j]);
            pot[i * n + j] = r <= R ? -alpha * r * r : alpha * R * R * (2 * std::log(R / r) - 1);
        }
    Dev dp(pot), dl((n - 1) * (n - 1));
    exterior::Grid<2> grid({c, c});
    exterior::laplacian<2>(nullptr, grid, index_last<2>(dl.p, {n - 1, n - 1}, 1), index_last<2>(dp.p, {n, n}, 1), 0);
    auto l = dl.host();
    for (int i = 1; i < n - 2; ++i) {
        double v = l[i * (n - 1) + n / 2];
        if (c[i] < -1.2 * R || c[i] > 1.2 * R) EXPECT_NEAR(v, 0., .5);
        else if (c[i] > -.8 * R && c[i] < .8 * R) EXPECT_NEAR(v, 1., .5);
    }
}

// tests/exterior/laplacian.cpp:630-732: staged and non-staged agree
TEST(Laplacian, Staged2D1Form)
{
    const int n = 31;
    auto c = uniform(-5., 5., n);
    const double R = 2.;
    std::vector<double> pot(n * n * 2);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double r = std::sqrt(c[i] * c[i] + c[j] * c[j]);
            double f = r <= R ? r / 3. - R / 2. : -R * R * R / (6. * r * r);
            pot[(i * n + j) * 2 + 0] = c[j] * f;
            pot[(i * n + j) * 2 + 1] = -c[i] * f;
        }
    Dev dp(pot), d1((n - 1) * (n - 1) * 2), d2((n - 1) * (n - 1) * 2);
    exterior::Grid<2> grid({c, c});
    auto in = index_last<2>(dp.p, {n, n}, 2);
    auto o1 = index_last<2>(d1.p, {n - 1, n - 1}, 2), o2 = index_last<2>(d2.p, {n - 1, n - 1}, 2);
    exterior::laplacian<2>(nullptr, grid, o1, in, 1);
    auto staged = exterior::make_staged_laplacian<2>(grid, o2, in, 1);
    staged.run(nullptr, o2, in);
    auto a = d1.host(), b = d2.host();
    for (size_t i = 0; i < a.size(); ++i) EXPECT_EQ(a[i], b[i]);
    for (int i = 1; i < n - 2; ++i) {
        double v = b[(i * (n - 1) + n / 2) * 2 + 1];
        if (c[i] < -1.2 * R || c[i] > 1.2 * R) EXPECT_NEAR(v, 0., .5);
        else if (c[i] > -.8 * R && c[i] < -.2 * R) EXPECT_NEAR(v, -1., .5);
        else if (c[i] > .2 * R && c[i] < .8 * R) EXPECT_NEAR(v, 1., .5);
    }
}

// tests/exterior/laplacian.cpp:745-835
TEST(Laplacian, 3D1Form)
{
    const int n = 21;
    auto c = uniform(-5., 5., n);
    const double R = 2., alpha = 0.2;
    std::vector<double> pot(n * n * n * 3);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < n; ++k) {
                double r = std::sqrt(c[i] * c[i] + c[j] * c[j] + c[k] * c[k]);
                double th = std::atan2(c[j], c[i]);
                double amp = r <= R ? alpha * r * r : alpha * R * R * R * R / (r * r);
                double* p = &pot[((i * n + j) * n + k) * 3];
                p[0] = amp * std::sin(th);
                p[1] = -amp * std::cos(th);
                p[2] = 0.;
            }
    const int m = n - 1;
    Dev dp(pot), dl(m * m * m * 3);
    exterior::Grid<3> grid({c, c, c});
    exterior::laplacian<3>(nullptr, grid, index_last<3>(dl.p, {m, m, m}, 3), index_last<3>(dp.p, {n, n, n}, 3), 1);
    auto l = dl.host();
    for (int i = 1; i < m - 1; ++i) {
        double v = l[((i * m + n / 2) * m + n / 2) * 3 + 1];
        if (c[i] < -1.3 * R || c[i] > 1.3 * R) EXPECT_NEAR(v, 0., .5);
        else if (c[i] > -.7 * R && c[i] < -.3 * R) EXPECT_NEAR(v, -1., .5);
        else if (c[i] > .3 * R && c[i] < .7 * R) EXPECT_NEAR(v, 1., .5);
    }
}

// minimize_strong_formulation_residual on the matrix-free operator and on its
// assembled matrix (use_matrix_free = true / false) with the Jacobi and SpdIsai
// preconditioners: every solve converges and all reach the same solution.
TEST(Solvers, AssembledMatchesMatrixFree)
{
    const int nx = 9, ny = 8, nz = 10;
    auto x = uniform(0., 1., nx), y = uniform(0., .8, ny), z = uniform(0., 1.2, nz);
    const int nn = nx * ny * nz;
    std::vector<double> mu(nn), rhs(3 * nn, 0.);
    for (int q = 0; q < nn; ++q) mu[q] = 1. + 0.5 * std::sin(0.37 * q);
    for (int k = 1; k < nz - 1; ++k)
        for (int j = 1; j < ny - 1; ++j)
            for (int i = 1; i < nx - 1; ++i)
                for (int c = 0; c < 3; ++c) rhs[3 * (i + nx * (j + ny * k)) + c] = std::cos(0.1 * (i + 2 * j + 3 * k + c));
    Dev dmu(mu), db(rhs), x0(3 * nn), x1(3 * nn), x2(3 * nn);
    sil_operator op = nullptr;
    if (sil_magnetostatics3d_create(&op, nx, ny, nz, x.data(), y.data(), z.data(), dmu.p, 20.0) != SIL_OK)
        throw std::runtime_error(sil_last_error());
    sil_operator m = solvers::build_matrix(nullptr, op);
    solvers::StrongFormulationSolverSettings st;
    st.relative_tolerance = 1e-11;
    auto d0 = solvers::minimize_strong_formulation_residual(nullptr, op, db.p, x0.p, st);
    auto d1 = solvers::minimize_strong_formulation_residual(nullptr, m, db.p, x1.p, st);
    st.preconditioner = SIL_PRECOND_SPD_ISAI;
    auto d2 = solvers::minimize_strong_formulation_residual(nullptr, m, db.p, x2.p, st);
    EXPECT_EQ(d0.converged, 1);
    EXPECT_EQ(d1.converged, 1);
    EXPECT_EQ(d2.converged, 1);
    EXPECT_NEAR(d1.iterations, d0.iterations, 2);
    EXPECT_EQ(d2.iterations < d1.iterations, 1);
    auto h0 = x0.host(), h1 = x1.host(), h2 = x2.host();
    double scale = 0., e1 = 0., e2 = 0.;
    for (size_t q = 0; q < h0.size(); ++q) {
        scale = std::fmax(scale, std::fabs(h0[q]));
        e1 = std::fmax(e1, std::fabs(h1[q] - h0[q]));
        e2 = std::fmax(e2, std::fabs(h2[q] - h0[q]));
    }
    EXPECT_NEAR(e1 / scale, 0., 1e-8);
    EXPECT_NEAR(e2 / scale, 0., 1e-8);
    sil_operator_destroy(m);
    sil_operator_destroy(op);
}

TEST(Solvers, Magnetostatics2DNewtonAndPostProcess)
{
    // MagnetostaticsOperator2D with a nonlinear block, solved by Newton
    // through the same solver entry point, then post-processed per cell
    const int nx = 20, ny = 18, nn = nx * ny, mx = nx - 1, my = ny - 1;
    std::vector<double> x(nx), y(ny), ferro(nn, 0.), mu(nn, 4e-7 * M_PI), jz(mx * my, 0.);
    for (int i = 0; i < nx; ++i) x[i] = i;
    for (int j = 0; j < ny; ++j) y[j] = j;
    for (int j = 6; j < 12; ++j)
        for (int i = 7; i < 13; ++i) ferro[i + nx * j] = 1.;
    for (int j = 7; j < 10; ++j)
        for (int i = 2; i < 5; ++i) {
            jz[i + mx * j] = 1e5;
            jz[(i + 12) + mx * j] = -1e5;
        }
    const double bh_b[] = {0., 0.5, 1.0, 1.5, 2.0}, bh_h[] = {0., 100., 400., 2000., 10000.};
    Dev dmu(mu), dfe(ferro), djz(jz), rhs(nn), sol(nn), cmu(std::vector<double>(mx * my, 4e-7 * M_PI));
    Dev B(3 * mx * my), H(3 * mx * my), T(6 * mx * my), F(3 * mx * my);
    sil_operator op = nullptr;
    if (sil_magnetostatics2d_nonlinear_create(&op, nx, ny, x.data(), y.data(), dmu.p, dfe.p, bh_b, bh_h, 5, 1) != SIL_OK)
        throw std::runtime_error(sil_last_error());
    EXPECT_EQ(sil_magnetostatics2d_rhs(op, djz.p, rhs.p, nullptr), SIL_OK);
    solvers::StrongFormulationSolverSettings st;
    st.relative_tolerance = 1e-10;
    st.max_iterations = 5000;
    auto d = solvers::minimize_strong_formulation_residual(nullptr, op, rhs.p, sol.p, st);
    EXPECT_EQ(d.converged, 1);
    EXPECT_EQ(d.outer_iterations >= 1, 1);
    // the nonlinear residual b − F(x), recomputed through the operator
    Dev fx(nn);
    EXPECT_EQ(sil_operator_apply(op, sol.p, fx.p, nullptr), SIL_OK);
    auto hb = rhs.host(), hf = fx.host();
    double rr = 0., bb = 0.;
    for (int q = 0; q < nn; ++q) {
        rr += (hb[q] - hf[q]) * (hb[q] - hf[q]);
        bb += hb[q] * hb[q];
    }
    EXPECT_NEAR(std::sqrt(rr / bb), 0., 1.0001e-10);
    std::vector<double> flags(mx * my, 0.);
    for (int j = 6; j < 11; ++j)
        for (int i = 7; i < 12; ++i) flags[i + mx * j] = 1.;
    Dev dfl(flags);
    EXPECT_EQ(sil_magnetostatics2d_post_process(op, sol.p, cmu.p, dfl.p, B.p, H.p, T.p, F.p, nullptr), SIL_OK);
    auto hB = B.host(), hT = T.host();
    double core = 0., air = 0.;
    for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
            const int c = i + mx * j;
            const double b = std::hypot(hB[3 * c], hB[3 * c + 1]);
            if (flags[c] != 0.) core = std::fmax(core, b);
            else if (i > 14 && j > 12) air = std::fmax(air, b);
            EXPECT_EQ(hB[3 * c + 2], 0.);
            EXPECT_EQ(std::isfinite(hT[6 * c + 2]), 1);
        }
    EXPECT_EQ(core > air, 1);  // the permeable block carries the flux
    sil_operator_destroy(op);
}

}  // namespace

int main()
{
    int failed_tests = 0;
    for (auto& [name, fn] : reg().tests) {
        int before = reg().failures;
        try {
            fn();
        } catch (std::exception const& e) {
            std::printf("  EXCEPTION %s\n", e.what());
            ++reg().failures;
        }
        bool ok = reg().failures == before;
        failed_tests += ok ? 0 : 1;
        std::printf("[%s] %s\n", ok ? "  OK  " : " FAIL ", name.c_str());
    }
    std::printf("%zu tests, %d checks, %d failed tests\n", reg().tests.size(), reg().checks, failed_tests);
    return failed_tests == 0 ? 0 : 1;
}
