// mmio.cpp -- Matrix Market coordinate reader and the load -> CSR -> lower
// triangle pipeline of the reference driver (main.cu:66-246), written from
// the Matrix Market format definition (NIST exchange format: a banner line
// "%%MatrixMarket <object> <format> <field> <symmetry>", '%' comment lines,
// a size line, then one entry per line with 1-based indices).
//
// API parity with header/mmio.h:25-52: mm_read_banner, mm_read_mtx_crd_size,
// mm_is_valid and the mm_is_* macros (in sptrsv_amd.h) keep their names and
// argument meaning; mm_typecode_to_str_r writes into a caller buffer.
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "sptrsv_amd.h"

namespace {

// Lower-case copy of a token.
std::string lower(const char* s) {
    std::string r(s);
    for (auto& c : r) c = (char)tolower((unsigned char)c);
    return r;
}

}  // namespace

namespace sptrsv {
int fail(int status, const std::string& msg);
}

namespace {
int set_err(const std::string& m) { return sptrsv::fail(SPTRSV_E_IO, m); }
}  // namespace

extern "C" {

// Status values follow the exchange format's reference codes in meaning:
// 0 ok, 11 read error, 12 premature end of file, 14 no header, 15
// unsupported type.
int mm_read_banner(FILE* f, MM_typecode* matcode) {
    if (!f || !matcode) return 11;
    char line[MM_MAX_LINE_LENGTH + 1];
    char b[MM_MAX_TOKEN_LENGTH], obj[MM_MAX_TOKEN_LENGTH], fmt[MM_MAX_TOKEN_LENGTH],
        fld[MM_MAX_TOKEN_LENGTH], sym[MM_MAX_TOKEN_LENGTH];
    (*matcode)[0] = (*matcode)[1] = (*matcode)[2] = ' ';
    (*matcode)[3] = 'G';
    if (!fgets(line, sizeof(line), f)) return 12;
    if (sscanf(line, "%63s %63s %63s %63s %63s", b, obj, fmt, fld, sym) != 5) return 12;
    if (strcmp(b, "%%MatrixMarket") != 0) return 14;
    std::string o = lower(obj), fm = lower(fmt), fi = lower(fld), sy = lower(sym);
    if (o != "matrix") return 15;
    (*matcode)[0] = 'M';
    if (fm == "coordinate") (*matcode)[1] = 'C';
    else if (fm == "array") (*matcode)[1] = 'A';
    else return 15;
    if (fi == "real") (*matcode)[2] = 'R';
    else if (fi == "complex") (*matcode)[2] = 'C';
    else if (fi == "pattern") (*matcode)[2] = 'P';
    else if (fi == "integer") (*matcode)[2] = 'I';
    else return 15;
    if (sy == "general") (*matcode)[3] = 'G';
    else if (sy == "symmetric") (*matcode)[3] = 'S';
    else if (sy == "hermitian") (*matcode)[3] = 'H';
    else if (sy == "skew-symmetric") (*matcode)[3] = 'K';
    else return 15;
    return 0;
}

int mm_read_mtx_crd_size(FILE* f, int* M, int* N, int* nz) {
    if (!f || !M || !N || !nz) return 11;
    char line[MM_MAX_LINE_LENGTH + 1];
    *M = *N = *nz = 0;
    do {
        if (!fgets(line, sizeof(line), f)) return 12;
    } while (line[0] == '%' || line[strspn(line, " \t\r\n")] == '\0');
    if (sscanf(line, "%d %d %d", M, N, nz) != 3) return 12;
    return 0;
}

int mm_is_valid(MM_typecode t) {
    if (!mm_is_matrix(t)) return 0;
    if (mm_is_dense(t) && mm_is_pattern(t)) return 0;
    if (mm_is_real(t) && mm_is_hermitian(t)) return 0;
    if (mm_is_pattern(t) && (mm_is_hermitian(t) || mm_is_skew(t))) return 0;
    return 1;
}

char* mm_typecode_to_str_r(const MM_typecode t, char* buf) {
    const char* f = mm_is_coordinate(t) ? "coordinate" : mm_is_array(t) ? "array" : "?";
    const char* v = mm_is_real(t) ? "real" : mm_is_complex(t) ? "complex"
                    : mm_is_pattern(t) ? "pattern" : mm_is_integer(t) ? "integer" : "?";
    const char* s = mm_is_general(t) ? "general" : mm_is_symmetric(t) ? "symmetric"
                    : mm_is_hermitian(t) ? "hermitian" : mm_is_skew(t) ? "skew-symmetric" : "?";
    snprintf(buf, 64, "%s %s %s %s", mm_is_matrix(t) ? "matrix" : "?", f, v, s);
    return buf;
}

void sptrsv_free_host(void* p) { free(p); }

int sptrsv_read_mtx_lower(const char* path, int unit, int* n_out, int* nnz_out, int** ia_out,
                          int** ja_out, double** a_out) {
    if (!path || !n_out || !nnz_out || !ia_out || !ja_out || !a_out)
        return sptrsv::fail(SPTRSV_E_INVALID_ARG, "null argument");
    *ia_out = nullptr;
    *ja_out = nullptr;
    *a_out = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f) return set_err(std::string("cannot open ") + path);
    MM_typecode t;
    int rc = mm_read_banner(f, &t);
    if (rc != 0) { fclose(f); return set_err("could not process the Matrix Market banner"); }
    if (mm_is_complex(t)) { fclose(f); return set_err("complex matrices are not supported"); }  // main.cu:83-86
    if (!mm_is_coordinate(t)) { fclose(f); return set_err("only coordinate format is supported"); }
    int m = 0, ncol = 0, nz = 0;
    if (mm_read_mtx_crd_size(f, &m, &ncol, &nz) != 0 || m < 0 || nz < 0) {
        fclose(f);
        return set_err("bad size line");
    }
    if (m != ncol) { fclose(f); return set_err("matrix is not square"); }  // main.cu:112-115
    const bool sym = mm_is_symmetric(t) || mm_is_hermitian(t);            // main.cu:117-122
    const bool pattern = mm_is_pattern(t);

    // Read the entries (main.cu:131-155), 1-based -> 0-based.
    std::vector<int> ri(nz), ci(nz);
    std::vector<double> v(nz);
    int got = 0;
    char line[MM_MAX_LINE_LENGTH + 1];
    while (got < nz && fgets(line, sizeof(line), f)) {
        const char* p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '%' || *p == '\n' || *p == '\r' || *p == '\0') continue;
        char* e;
        long r = strtol(p, &e, 10);
        if (e == p) break;
        p = e;
        long c = strtol(p, &e, 10);
        if (e == p) break;
        p = e;
        double val = 1.0;
        if (!pattern) {
            val = strtod(p, &e);
            if (e == p) break;
        }
        if (r < 1 || r > m || c < 1 || c > m) {
            fclose(f);
            return set_err("entry index out of range at entry " + std::to_string(got + 1));
        }
        ri[got] = (int)r - 1;
        ci[got] = (int)c - 1;
        v[got] = val;
        ++got;
    }
    fclose(f);
    if (got != nz) return set_err("premature end of file: read " + std::to_string(got) + " of " +
                                  std::to_string(nz) + " entries");

    // COO -> CSR in file order with symmetric mirroring (main.cu:124-212).
    std::vector<int> cnt(m + 1, 0);
    for (int i = 0; i < nz; ++i) {
        cnt[ri[i]]++;
        if (sym && ri[i] != ci[i]) cnt[ci[i]]++;
    }
    std::vector<long long> rp(m + 1, 0);
    for (int i = 0; i < m; ++i) rp[i + 1] = rp[i] + cnt[i];
    long long nnzA = rp[m];
    std::vector<int> col(nnzA);
    std::vector<double> val(nnzA);
    std::fill(cnt.begin(), cnt.end(), 0);
    for (int i = 0; i < nz; ++i) {
        int r = ri[i], c = ci[i];
        long long o = rp[r] + cnt[r]++;
        col[o] = c;
        val[o] = v[i];
        if (sym && r != c) {
            o = rp[c] + cnt[c]++;
            col[o] = r;
            val[o] = v[i];
        }
    }
    // Lower triangle with explicit diagonal (main.cu:216-246): per row, copy
    // entries with col < i in storage order up to the first col >= i, then
    // append the diagonal.  unit: all values 1.0 (main.cu:228,236).
    long long cap = nnzA + m;
    int* ia = (int*)malloc(sizeof(int) * (size_t)(m + 1));
    int* ja = (int*)malloc(sizeof(int) * (size_t)(cap > 0 ? cap : 1));
    double* a = (double*)malloc(sizeof(double) * (size_t)(cap > 0 ? cap : 1));
    if (!ia || !ja || !a) {
        free(ia); free(ja); free(a);
        return sptrsv::fail(SPTRSV_E_ALLOC, "host allocation");
    }
    long long p = 0;
    ia[0] = 0;
    for (int i = 0; i < m; ++i) {
        double diag = 1.0;
        if (!unit)
            for (long long j = rp[i]; j < rp[i + 1]; ++j)
                if (col[j] == i) { diag = val[j]; break; }
        for (long long j = rp[i]; j < rp[i + 1]; ++j) {
            if (col[j] < i) {
                ja[p] = col[j];
                a[p] = unit ? 1.0 : val[j];
                ++p;
            } else {
                break;
            }
        }
        ja[p] = i;
        a[p] = diag;
        ++p;
        if (p > 0x7fffffffLL) {
            free(ia); free(ja); free(a);
            return sptrsv::fail(SPTRSV_E_CAPACITY, "lower triangle exceeds int32 nnz");
        }
        ia[i + 1] = (int)p;
    }
    *n_out = m;
    *nnz_out = (int)p;
    *ia_out = ia;
    *ja_out = ja;
    *a_out = a;
    return SPTRSV_OK;
}

}  // extern "C"
