// ref_matrices.cpp -- C entry points over the reference's cuda_test/matrices.h
// (SMS loader, CSR transpose, row/column rescaling) and cuda_test/wdmfiles.h
// (the .wdm checkpoint writer), compiled from the
// reference checkout by oracle/Makefile into oracle/_ref/libref.so.  This is a
// separate translation unit because matrices.h and sigma_basis.h both define
// display_vector (matrices.h:392, sigma_basis.h:31).
//
// TEST INFRASTRUCTURE ONLY (golden-vector generation and CPU cross-checks).
//
// T = int64_t: CooMatrix::from_sms_file (matrices.h:241-280) maps a negative
// value v to ((v % p) + p) % p, which is only right for a signed T.
#include <cstdint>
#include <string>
#include "matrices.h"
#include "wdmfiles.h"

extern "C" {

// Loads an SMS file.  First call with col/val NULL returns m, n, nnz; second call fills.
int ref_load_sms(const char *path, uint32_t p, int64_t *m, int64_t *n, int64_t *nnz,
                 int64_t *row_ptr, int32_t *col, uint32_t *val) {
    CsrMatrix<int64_t> A = CsrMatrix<int64_t>::from_sms_file(std::string(path), (int64_t)p);
    *m = A.numRows;
    *n = A.numCols;
    *nnz = (int64_t)A.values.size();
    if (!row_ptr) return 0;
    for (int i = 0; i <= A.numRows; i++) row_ptr[i] = A.rowOffsets[i];
    for (size_t e = 0; e < A.values.size(); e++) { col[e] = A.colIndices[e]; val[e] = (uint32_t)A.values[e]; }
    return 0;
}

static CsrMatrix<int64_t> mk(int64_t m, int64_t n, const int64_t *rp, const int32_t *col, const uint32_t *val) {
    std::vector<int> off(rp, rp + m + 1), idx(col, col + rp[m]);
    std::vector<int64_t> v(val, val + rp[m]);
    return CsrMatrix<int64_t>((int)m, (int)n, off, idx, v);
}

// matrices.h:94-149
int ref_transpose(int64_t m, int64_t n, const int64_t *rp, const int32_t *col, const uint32_t *val,
                  int64_t *rpt, int32_t *colt, uint32_t *valt) {
    CsrMatrix<int64_t> A = mk(m, n, rp, col, val);
    CsrMatrix<int64_t> T = A.transpose();
    for (int i = 0; i <= T.numRows; i++) rpt[i] = T.rowOffsets[i];
    for (size_t e = 0; e < T.values.size(); e++) { colt[e] = T.colIndices[e]; valt[e] = (uint32_t)T.values[e]; }
    return 0;
}

// matrices.h:25-63: values <- values * d_col mod p, then (if dr) * d_row mod p
int ref_rescale(int64_t m, int64_t n, const int64_t *rp, const int32_t *col, uint32_t *val,
                const uint32_t *dr, const uint32_t *dc, uint32_t p) {
    CsrMatrix<int64_t> A = mk(m, n, rp, col, val);
    if (dr) {
        std::vector<int64_t> s(dr, dr + m);
        A.csr_rowrescale(s, (int64_t)p);
    }
    if (dc) {
        std::vector<int64_t> s(dc, dc + n);
        A.csr_columnrescale(s, (int64_t)p);
    }
    for (size_t e = 0; e < A.values.size(); e++) val[e] = (uint32_t)A.values[e];
    return 0;
}

// wdmfiles.h:92-160: V and curv as b column lists of n values, seq as P pair
// lists of len values (pairs i <= j row-major).
int ref_save_wdm(const char *path, int64_t m, int64_t n, uint32_t p, uint32_t b, const uint32_t *dr,
                 const uint32_t *dc, const uint32_t *V, const uint32_t *curv, const uint32_t *seq, int64_t len) {
    std::vector<uint32_t> r(dr, dr + m), c(dc, dc + n);
    std::vector<std::vector<uint32_t>> v(b, std::vector<uint32_t>(n)), cv(b, std::vector<uint32_t>(n));
    for (uint32_t i = 0; i < b; i++)
        for (int64_t k = 0; k < n; k++) {
            v[i][k] = V[k * b + i];
            cv[i][k] = curv[k * b + i];
        }
    const uint32_t P = b * (b + 1) / 2;
    std::vector<std::vector<uint32_t>> s(P, std::vector<uint32_t>(len));
    for (uint32_t ii = 0; ii < P; ii++)
        for (int64_t t = 0; t < len; t++) s[ii][t] = seq[t * P + ii];
    save_wdm_file_sym<uint32_t>(std::string(path), (size_t)m, (size_t)n, p, r, c, v, cv, s);
    return 0;
}

}  // extern "C"
