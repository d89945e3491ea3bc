// ref_algebra.cpp -- thin C entry points over the reference's own CPU C++
// algebra headers (cuda_test/ntt.h, poly_mat_mul.h, modular_linalg.h,
// sigma_basis.h), compiled from where they lie under the reference checkout by
// oracle/Makefile into oracle/_ref/libref.so.
//
// TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py to write
// golden vectors and by the CPU tests to cross-check the oracle.  Nothing on
// the product path loads it.
//
// <iomanip> is included first because cuda_test/sigma_basis.h:82 uses
// std::setprecision with its own #include <iomanip> commented out (:8).
#include <iomanip>
#include <sstream>
#include <string>
#include <cstdint>
#include "sigma_basis.h"

typedef std::vector<std::vector<std::vector<uint64_t>>> PM64;

extern "C" {

// cuda_test/ntt.h:152-193, ntt<uint64_t>: NTT mod 2^61 - 2^26 + 1, root 3.
// (ntt<uint32_t> does not instantiate: mod_pow(root, p - 1 - (p - 1) / n) at
// ntt.h:170-171 deduces conflicting types uint32_t / size_t.)
int ref_ntt_u64(uint64_t *a, uint64_t n, int invert) {
    std::vector<uint64_t> v(a, a + n);
    ntt(v, invert != 0);
    std::copy(v.begin(), v.end(), a);
    return 0;
}

// cuda_test/poly_mat_mul.h:227-251, poly_mat_mul_fft_red<uint64_t>: product lifted to
// 2^61 - 2^26 + 1 and reduced mod reducetoprime (exact while the integer result fits).
// A [m][n][na], B [n][k][nb] -> C [m][k][na + nbb - 1], nbb = min(nb, b_end - b_start).
int ref_poly_mat_mul_u64(uint32_t m, uint32_t n, uint32_t k, const uint32_t *A, uint64_t na,
                         const uint32_t *B, uint64_t nb, uint64_t b_start, uint64_t b_end,
                         uint32_t reducetoprime, uint32_t *C) {
    std::vector<std::vector<std::vector<uint64_t>>> a(m, std::vector<std::vector<uint64_t>>(n)),
        b(n, std::vector<std::vector<uint64_t>>(k));
    for (uint32_t i = 0; i < m; i++)
        for (uint32_t j = 0; j < n; j++) a[i][j].assign(A + ((size_t)i * n + j) * na, A + ((size_t)i * n + j + 1) * na);
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < k; j++) b[i][j].assign(B + ((size_t)i * k + j) * nb, B + ((size_t)i * k + j + 1) * nb);
    auto c = poly_mat_mul_fft_red<uint64_t>(a, b, (uint64_t)reducetoprime, b_start, b_end);
    size_t len = c[0][0].size();
    for (uint32_t i = 0; i < m; i++)
        for (uint32_t j = 0; j < k; j++)
            for (size_t l = 0; l < len; l++) C[((size_t)i * k + j) * len + l] = (uint32_t)c[i][j][l];
    return (int)len;
}

// cuda_test/modular_linalg.h:287-359 with T = uint64_t.
int ref_lsp_u64(uint32_t m, const uint32_t *a, uint32_t p, uint32_t *L, uint32_t *S, uint32_t *P) {
    DMatrix<uint64_t> A(m, m);
    for (size_t i = 0; i < (size_t)m * m; i++) A.data[i] = a[i];
    auto [l, s, pm] = lsp_decomposition(A, (uint64_t)p);
    for (size_t i = 0; i < (size_t)m * m; i++) { L[i] = (uint32_t)l.data[i]; S[i] = (uint32_t)s.data[i]; P[i] = (uint32_t)pm.data[i]; }
    return 0;
}

// cuda_test/modular_linalg.h:107-170 with T = uint64_t.  Returns -1 when singular.
int ref_inverse_u64(uint32_t n, const uint32_t *a, uint32_t p, uint32_t *out) {
    DMatrix<uint64_t> A(n, n);
    for (size_t i = 0; i < (size_t)n * n; i++) A.data[i] = a[i];
    try {
        auto inv = matrix_inverse(A, (uint64_t)p);
        for (size_t i = 0; i < (size_t)n * n; i++) out[i] = (uint32_t)inv.data[i];
    } catch (const std::exception &) {
        return -1;
    }
    return 0;
}

// cuda_test/modular_linalg.h:234-285 with T = uint64_t.
uint64_t ref_modular_rank_u64(uint32_t rows, uint32_t cols, const uint32_t *a, uint32_t p) {
    DMatrix<uint64_t> A(rows, cols);
    for (size_t i = 0; i < (size_t)rows * cols; i++) A.data[i] = a[i];
    return modular_rank(A, (uint64_t)p);
}

// cuda_test/sigma_basis.h:91-171, M_Basis<uint64_t> on a constant m x n matrix.
int ref_m_basis_u64(uint32_t m, uint32_t n, const uint32_t *G0, int64_t *delta, uint32_t p, uint32_t *M) {
    PM64 G(m, std::vector<std::vector<uint64_t>>(n, std::vector<uint64_t>(1)));
    for (uint32_t i = 0; i < m; i++)
        for (uint32_t j = 0; j < n; j++) G[i][j][0] = G0[(size_t)i * n + j];
    std::vector<int64_t> del(delta, delta + m);
    auto [R, d2] = M_Basis<uint64_t>(G, del, (uint64_t)p);
    for (uint32_t i = 0; i < m; i++)
        for (uint32_t j = 0; j < m; j++) {
            M[((size_t)i * m + j) * 2] = (uint32_t)R[i][j][0];
            M[((size_t)i * m + j) * 2 + 1] = (uint32_t)R[i][j][1];
        }
    std::copy(d2.begin(), d2.end(), delta);
    return 0;
}

// cuda_test/sigma_basis.h:214-263, PM_Basis<uint64_t, uint64_t>.  The products lift to
// the 61-bit prime, so this is exact only while (d+1)*2b*p^2 < 2^61 (small p).
// seq [L][b(b+1)/2] t-major.  d is rounded up to a power of two as the reference does.
// basis_out [2b][2b][len] with len returned.
int64_t ref_pm_basis_u64(const uint32_t *seq, uint32_t b, uint64_t L, uint64_t d, uint32_t p,
                         uint32_t *basis_out, uint64_t cap, int64_t *delta_out) {
    uint32_t P = b * (b + 1) / 2;
    std::vector<std::vector<uint64_t>> s(P, std::vector<uint64_t>(L));
    for (uint64_t t = 0; t < L; t++)
        for (uint32_t ii = 0; ii < P; ii++) s[ii][t] = seq[t * P + ii];
    std::ostringstream sink;
    auto *old = std::cout.rdbuf(sink.rdbuf());
    auto [M, mu] = PM_Basis<uint64_t, uint64_t>(s, d, (uint64_t)p);
    std::cout.rdbuf(old);
    size_t n2 = 2 * b, len = M[0][0].size();
    if ((uint64_t)n2 * n2 * len > cap) return -1;
    for (size_t i = 0; i < n2; i++)
        for (size_t j = 0; j < n2; j++)
            for (size_t k = 0; k < len; k++) basis_out[(i * n2 + j) * len + k] = (uint32_t)M[i][j][k];
    for (size_t i = 0; i < n2; i++) delta_out[i] = mu[i];
    return (int64_t)len;
}

// cuda_test/sigma_basis.h:265-360, analyze_pm_basis<uint64_t>.  It prints the rank
// estimate; the line "Estimated matrix rank: N" is captured and parsed.
int64_t ref_analyze_pm_basis_u64(const uint32_t *basis, const int64_t *delta, uint32_t b, uint64_t len, uint32_t p) {
    size_t n2 = 2 * b;
    PM64 B(n2, std::vector<std::vector<uint64_t>>(n2, std::vector<uint64_t>(len)));
    for (size_t i = 0; i < n2; i++)
        for (size_t j = 0; j < n2; j++)
            for (size_t k = 0; k < len; k++) B[i][j][k] = basis[(i * n2 + j) * len + k];
    std::vector<int64_t> del(delta, delta + n2);
    std::ostringstream sink;
    auto *old = std::cout.rdbuf(sink.rdbuf());
    analyze_pm_basis<uint64_t>(B, del, (uint64_t)p);
    std::cout.rdbuf(old);
    std::string out = sink.str();
    const std::string key = "Estimated matrix rank: ";
    size_t pos = out.rfind(key);
    if (pos == std::string::npos) return -1;
    return std::stoll(out.substr(pos + key.size()));
}

}  // extern "C"
