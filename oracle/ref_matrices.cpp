// ref_matrices.cpp -- C entry points over the reference's cuda_test/matrices.h
// (SMS loader, CSR transpose, row/column rescaling), compiled from the
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

}  // extern "C"
