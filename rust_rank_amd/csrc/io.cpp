// io.cpp -- the text formats around the reference's rank pipeline:
//   - the SMS sparse-matrix loader, src/matrices.rs:662-732
//     (CsrMatrix::load_csr_matrix_from_sms);
//   - the .wdm checkpoint files, src/wdm_files.rs:10-210
//     (load_wdm_file_sym / save_wdm_file_sym);
//   - the number of entries the sequence loop appends on a resume,
//     src/main.rs:105-131.
// Host code only.  Parsing follows the Rust loaders token by token: tokens
// are split on whitespace, unsigned fields accept an optional '+', signed
// ones a sign, and anything else (or an out-of-range value) is an error.
// Writing uses the reference's layout: one header line, then one line per
// vector, values separated by single spaces.
#include "common.hpp"
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace rr {

namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// One line of a text file at a time (lines of any length).
struct LineReader {
    FILE *f = nullptr;
    char *line = nullptr;
    size_t cap = 0;
    ssize_t len = 0;
    bool open(const char *path) { return (f = fopen(path, "rb")) != nullptr; }
    // false at end of file (the line is then empty, as Rust's read_line leaves it)
    bool next() {
        len = f ? getline(&line, &cap, f) : -1;
        if (len < 0) {
            len = 0;
            return false;
        }
        return true;
    }
    ~LineReader() {
        free(line);
        if (f) fclose(f);
    }
};

// Whitespace-separated tokens of [b, e).
struct Tokens {
    const char *p, *e;
    Tokens(const char *b, const char *end) : p(b), e(end) {}
    bool next(const char *&tb, const char *&te) {
        while (p < e && is_space(*p)) p++;
        if (p == e) return false;
        tb = p;
        while (p < e && !is_space(*p)) p++;
        te = p;
        return true;
    }
};

// u64 of [b, e): optional '+', decimal digits, at most max.
bool parse_unsigned(const char *b, const char *e, uint64_t max, uint64_t &out) {
    if (b < e && *b == '+') b++;
    if (b == e) return false;
    uint64_t v = 0;
    for (; b < e; b++) {
        if (*b < '0' || *b > '9') return false;
        const uint64_t d = (uint64_t)(*b - '0');
        if (v > (max - d) / 10) return false;
        v = v * 10 + d;
    }
    out = v;
    return true;
}

// i32 of [b, e): optional sign, decimal digits, in [-2^31, 2^31).
bool parse_i32(const char *b, const char *e, int64_t &out) {
    bool neg = false;
    if (b < e && (*b == '+' || *b == '-')) neg = *b++ == '-';
    uint64_t v;
    if (!parse_unsigned(b, e, neg ? (1ull << 31) : (1ull << 31) - 1, v)) return false;
    out = neg ? -(int64_t)v : (int64_t)v;
    return true;
}

// Decimal writer into a large buffer.
struct Writer {
    FILE *f = nullptr;
    std::vector<char> buf;
    size_t n = 0;
    bool ok = true;
    explicit Writer(FILE *f_) : f(f_), buf(1 << 22) {}
    void flush() {
        if (n && fwrite(buf.data(), 1, n, f) != n) ok = false;
        n = 0;
    }
    void room(size_t k) {
        if (n + k > buf.size()) flush();
    }
    void ch(char c) {
        room(1);
        buf[n++] = c;
    }
    void num(uint64_t v) {
        char t[24];
        int k = 0;
        do {
            t[k++] = (char)('0' + v % 10);
            v /= 10;
        } while (v);
        room((size_t)k);
        while (k) buf[n++] = t[--k];
    }
};

}  // namespace

}  // namespace rr

// Matrix read by rr_sms_open, held until rr_sms_close.
struct rr_sms {
    uint64_t m = 0, n = 0;
    std::vector<uint64_t> row_ptr;
    std::vector<uint32_t> col, val;
};

using namespace rr;

extern "C" {

rr_status rr_sms_open(const char *path, uint32_t p, rr_sms **out, uint64_t *m, uint64_t *n, uint64_t *nnz) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(path && out && m && n && nnz, RR_EINVAL, "null argument");
        RR_CHECK(p >= 2 && p < (1u << 31), RR_EINVAL, "p=%u must be in [2, 2^31)", p);
        *out = nullptr;
        LineReader R;
        RR_CHECK(R.open(path), RR_EINVAL, "cannot open %s", path);
        std::unique_ptr<rr_sms> S(new rr_sms());
        S->row_ptr.push_back(0);
        uint64_t lineno = 0;
        try {
            while (R.next()) {
                lineno++;
                const char *b = R.line, *e = R.line + R.len;
                while (b < e && is_space(*b)) b++;
                while (e > b && is_space(e[-1])) e--;
                if (b == e || *b == '%') continue;  // comments and empty lines
                const char *tb[4], *te[4];
                int k = 0;
                Tokens T(b, e);
                while (k < 4 && T.next(tb[k], te[k])) k++;
                if (k != 3) continue;  // only three-token lines are read
                if (S->m == 0 && S->n == 0) {  // the dimensions (the third field is not used)
                    uint64_t r, c;
                    RR_CHECK(parse_unsigned(tb[0], te[0], ~0ull, r) && parse_unsigned(tb[1], te[1], ~0ull, c),
                             RR_EINVAL, "%s:%llu: bad dimensions", path, (unsigned long long)lineno);
                    S->m = r;
                    S->n = c;
                    continue;
                }
                uint64_t row, col;
                int64_t v;
                RR_CHECK(parse_unsigned(tb[0], te[0], ~0ull, row) && parse_unsigned(tb[1], te[1], ~0ull, col) &&
                             parse_i32(tb[2], te[2], v),
                         RR_EINVAL, "%s:%llu: bad entry", path, (unsigned long long)lineno);
                int64_t r = v % (int64_t)p;  // remainder of the signed value, then into [0, p)
                if (r < 0) r += p;
                if (row == 0 || col == 0) continue;  // 1-based: zero indices (the "0 0 0" end line) are skipped
                RR_CHECK(row <= S->m && col <= S->n, RR_EINVAL, "%s:%llu: entry (%llu, %llu) outside %llux%llu", path,
                         (unsigned long long)lineno, (unsigned long long)row, (unsigned long long)col,
                         (unsigned long long)S->m, (unsigned long long)S->n);
                // rows open as the entries reach them; an entry of an earlier row joins the open row, as in
                // the reference
                while (S->row_ptr.size() < row) S->row_ptr.push_back(S->val.size());
                S->val.push_back((uint32_t)r);
                S->col.push_back((uint32_t)(col - 1));
            }
            while (S->row_ptr.size() <= S->m) S->row_ptr.push_back(S->val.size());
        } catch (const std::bad_alloc &) {
            set_error("host allocation failed reading %s", path);
            return RR_ENOMEM;
        }
        RR_CHECK(S->m < (1ull << 31) && S->n < (1ull << 31) && S->val.size() < (1ull << 32), RR_EINVAL,
                 "%s: matrix too large (m, n < 2^31, nnz < 2^32)", path);
        *m = S->m;
        *n = S->n;
        *nnz = S->val.size();
        *out = S.release();
        return RR_OK;
    });
}

rr_status rr_sms_copy(const rr_sms *h, uint64_t *row_ptr, uint32_t *col, uint32_t *val) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(h && row_ptr && (h->val.empty() || (col && val)), RR_EINVAL, "null argument");
        std::memcpy(row_ptr, h->row_ptr.data(), (h->m + 1) * sizeof(uint64_t));
        if (!h->val.empty()) {
            std::memcpy(col, h->col.data(), h->col.size() * 4);
            std::memcpy(val, h->val.data(), h->val.size() * 4);
        }
        return RR_OK;
    });
}

void rr_sms_close(rr_sms *h) { delete h; }

rr_status rr_wdm_info(const char *path, uint64_t *m, uint64_t *n, uint32_t *p, uint64_t *len, uint32_t *b) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(path && m && n && p && len && b, RR_EINVAL, "null argument");
        LineReader R;
        RR_CHECK(R.open(path), RR_EINVAL, "cannot open %s", path);
        R.next();
        Tokens T(R.line, R.line + R.len);
        const char *names[5] = {"m", "n", "p", "Nlen", "num_v"};
        uint64_t v[5];
        for (int k = 0; k < 5; k++) {
            const char *tb, *te;
            RR_CHECK(T.next(tb, te), RR_EINVAL, "%s: Missing %s", path, names[k]);
            RR_CHECK(parse_unsigned(tb, te, k == 2 ? 0xFFFFFFFFull : ~0ull, v[k]), RR_EINVAL, "%s: bad %s", path,
                     names[k]);
        }
        RR_CHECK(v[4] >= 1 && v[4] <= 0xFFFFFFFFull, RR_EINVAL, "%s: num_v=%llu", path, (unsigned long long)v[4]);
        *m = v[0];
        *n = v[1];
        *p = (uint32_t)v[2];
        *len = v[3];
        *b = (uint32_t)v[4];
        return RR_OK;
    });
}

rr_status rr_wdm_load(const char *path, uint32_t *row_precond, uint32_t *col_precond, uint32_t *V, uint32_t *curv,
                      uint32_t *seq) {
    return rr::guarded([&]() -> rr_status {
        uint64_t m, n, len;
        uint32_t p, b;
        rr_status s = rr_wdm_info(path, &m, &n, &p, &len, &b);
        if (s != RR_OK) return s;
        RR_CHECK(row_precond && col_precond && V && curv && (seq || len == 0), RR_EINVAL, "null argument");
        LineReader R;
        RR_CHECK(R.open(path), RR_EINVAL, "cannot open %s", path);
        R.next();  // the header
        const uint64_t P = (uint64_t)b * (b + 1) / 2;
        // one line of unsigned values: dst[k * stride] for the first `take` tokens;
        // returns the number of tokens (all parsed when take is ~0)
        auto line = [&](uint32_t *dst, uint64_t stride, uint64_t cap, uint64_t take, uint64_t &count) -> bool {
            R.next();
            Tokens T(R.line, R.line + R.len);
            const char *tb, *te;
            count = 0;
            while (count < take && T.next(tb, te)) {
                uint64_t x;
                if (!parse_unsigned(tb, te, 0xFFFFFFFFull, x)) return false;
                if (count < cap) dst[count * stride] = (uint32_t)x;
                count++;
            }
            return true;
        };
        uint64_t cnt;
        RR_CHECK(line(row_precond, 1, m, ~0ull, cnt), RR_EINVAL, "%s: bad row preconditioner", path);
        RR_CHECK(cnt == m, RR_EINVAL, "%s: Row preconditioner length does not match matrix rows", path);
        RR_CHECK(line(col_precond, 1, n, ~0ull, cnt), RR_EINVAL, "%s: bad column preconditioner", path);
        RR_CHECK(cnt == n, RR_EINVAL, "%s: Column preconditioner length does not match matrix columns", path);
        for (uint32_t i = 0; i < b; i++) {
            RR_CHECK(line(V + i, b, n, ~0ull, cnt), RR_EINVAL, "%s: bad v vector %u", path, i);
            RR_CHECK(cnt == n, RR_EINVAL, "%s: v vector length does not match matrix columns", path);
        }
        for (uint32_t i = 0; i < b; i++) {
            RR_CHECK(line(curv + i, b, n, ~0ull, cnt), RR_EINVAL, "%s: bad curv vector %u", path, i);
            RR_CHECK(cnt == n, RR_EINVAL, "%s: curv vector length does not match matrix columns", path);
        }
        for (uint64_t ii = 0; ii < P; ii++) {  // pair ii's values over t, into the t-major [len][P] layout
            RR_CHECK(line(seq + ii, P, len, len, cnt), RR_EINVAL, "%s: bad sequence line %llu", path,
                     (unsigned long long)ii);
            RR_CHECK(cnt == len, RR_EINVAL, "%s: seq vector length does not match expected sequence length", path);
        }
        return RR_OK;
    });
}

rr_status rr_wdm_save(const char *path, uint64_t m, uint64_t n, uint32_t p, uint32_t b, const uint32_t *row_precond,
                      const uint32_t *col_precond, const uint32_t *V, const uint32_t *curv, const uint32_t *seq,
                      uint64_t len) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(path && row_precond && col_precond && V && curv && (seq || len == 0), RR_EINVAL, "null argument");
        RR_CHECK(b >= 1 && p >= 2, RR_EINVAL, "b=%u, p=%u", b, p);
        FILE *f = fopen(path, "wb");
        RR_CHECK(f, RR_EINVAL, "cannot create %s", path);
        Writer W(f);
        const uint64_t P = (uint64_t)b * (b + 1) / 2;
        W.num(m), W.ch(' '), W.num(n), W.ch(' '), W.num(p), W.ch(' '), W.num(len), W.ch(' '), W.num(b), W.ch('\n');
        auto row = [&](const uint32_t *x, uint64_t count, uint64_t stride, bool canon) {
            for (uint64_t k = 0; k < count; k++) {
                if (k) W.ch(' ');
                const uint64_t v = x[k * stride];
                W.num(canon ? (v + p) % p : v);  // prettify_vect (src/matrices.rs:67-71) on v and curv
            }
            W.ch('\n');
        };
        row(row_precond, m, 1, false);
        row(col_precond, n, 1, false);
        for (uint32_t i = 0; i < b; i++) row(V + i, n, b, true);
        for (uint32_t i = 0; i < b; i++) row(curv + i, n, b, true);
        for (uint64_t ii = 0; ii < P; ii++) row(seq + ii, len, P, false);
        W.flush();
        const bool ok = W.ok && fclose(f) == 0;
        RR_CHECK(ok, RR_EINVAL, "writing %s failed", path);
        return RR_OK;
    });
}

uint64_t rr_seq_resume_count(uint64_t have, uint64_t max_nlen) {
    if (have > max_nlen) return 0;
    const uint64_t pairs = (max_nlen - have) / 2;  // to_be_produced / 2 loop iterations, two entries each
    return (have == 0 ? 1 : 0) + 2 * pairs;        // a fresh sequence first gets S_0
}

}  // extern "C"
