// sdca_host.cpp -- host-side alignment pipeline and coupling writer of libsdca.so.
//
// The reference delegates parsing, filtering and the 4-state transform to apegrunt
// (not vendored; call sites SuperDCA/src/SuperDCA.cpp:173,274-277,354).  Its documented
// contract (README.md:32: loci with more than one non-gap allele, minor-allele frequency
// at least 1 %, gap frequency at most 15 %; a 4-state envelope of three alleles plus
// gap) is implemented here with these stated choices: A/C/G/T (any case) are alleles,
// every other character is a gap; the minor-allele frequency is the count of the second
// most frequent allele over the non-gap count; alleles are numbered by descending count
// with ties broken A<C<G<T; the gap is the last envelope state; loci with more alleles
// than the envelope holds are dropped.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sdca.h"
#include "sdca_format.hpp"

int sdca_fail(int code, const char *fmt, ...);  // sets the message of sdca_last_error (sdca_api.cpp)

namespace sdca {

// Offset of J_r[j](a, s) in SuperDCA's parameter vector: 16-locus blocks of q*q*16
// values, inside a block STATES access order (state*extent + position)*q + s, the last
// block having extent L % 16 (SuperDCA/include/Matrix_kernel_access_order.hpp:28-34,
// SuperDCA/include/Coupling_matrix_view.hpp:137-160).  h_r(s) follows at L*q*q + s.
size_t ref_index(int L, int q, int j, int a, int s)
{
    const size_t blk = (size_t)j / 16, pos = (size_t)j % 16;
    const size_t ext = ((blk + 1) * 16 > (size_t)L) ? (size_t)L % 16 : 16;
    return blk * 16 * (size_t)q * q + ((size_t)a * ext + pos) * q + (size_t)s;
}

}  // namespace sdca

namespace {
int nuc_index(char ch)
{
    switch (ch) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return -1;
    }
}

}  // namespace

using sdca::put_fixed8;
using sdca::put_uint;

extern "C" {

int sdca_prepare_alignment(const char *chars, int32_t B, int32_t L_in, double maf_threshold, double gap_threshold,
                           int32_t no_filter, int32_t q_env, uint8_t *states_out, int32_t *loci_out, int32_t *L_out)
{
    if (!chars || !states_out || !loci_out || !L_out || B < 1 || L_in < 0 || q_env < 2 || q_env > 8)
        return sdca_fail(SDCA_E_ARG, "sdca_prepare_alignment: null buffer, B < 1, L_in < 0 or q_env outside [2, 8]");
    std::vector<int32_t> kept;
    std::vector<std::array<int, 4>> maps;
    kept.reserve(L_in);
    for (int j = 0; j < L_in; ++j) {
        int cnt[4] = {0, 0, 0, 0};
        int ngap = 0;
        for (int b = 0; b < B; ++b) {
            const int a = nuc_index(chars[(size_t)b * L_in + j]);
            if (a < 0) ++ngap;
            else ++cnt[a];
        }
        int order[4] = {0, 1, 2, 3};
        // stable descending sort by count (ties keep A<C<G<T)
        for (int u = 1; u < 4; ++u)
            for (int v = u; v > 0 && cnt[order[v]] > cnt[order[v - 1]]; --v) std::swap(order[v], order[v - 1]);
        int nall = 0;
        for (int u = 0; u < 4; ++u) nall += cnt[u] > 0;
        const int nongap = B - ngap;
        const double maf = (nall > 1 && nongap > 0) ? (double)cnt[order[1]] / (double)nongap : 0.0;
        const double gapf = (double)ngap / (double)B;
        const bool pass = no_filter || (nall > 1 && maf >= maf_threshold && gapf <= gap_threshold);
        if (!pass || nall > q_env - 1) continue;
        std::array<int, 4> mp;
        for (int u = 0; u < 4; ++u) mp[order[u]] = u;
        kept.push_back(j);
        maps.push_back(mp);
    }
    const int L = (int)kept.size();
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < L; ++k) {
            const int a = nuc_index(chars[(size_t)b * L_in + kept[k]]);
            states_out[(size_t)b * L + k] = (uint8_t)(a < 0 ? q_env - 1 : maps[k][a]);
        }
    for (int k = 0; k < L; ++k) loci_out[k] = kept[k];
    *L_out = L;
    return SDCA_OK;
}

int sdca_write_couplings(const char *path, const float *S, int32_t L, const int32_t *translation, int32_t base)
{
    if (!path || !S || !translation || L < 0)
        return sdca_fail(SDCA_E_ARG, "sdca_write_couplings: null argument or L < 0");
    FILE *fp = fopen(path, "wb");
    if (!fp) return sdca_fail(SDCA_E_IO, "cannot open \"%s\" for writing", path);
    std::vector<char> buf(1 << 22);
    char *p = buf.data();
    char *const end = buf.data() + buf.size() - 128;
    int rc = SDCA_OK;
    for (int r = 0; r < L && rc == SDCA_OK; ++r) {
        const uint64_t ri = (uint64_t)((int64_t)translation[r] + base);
        for (int n = 0; n < r; ++n) {
            const float jij = S[(size_t)r * L + n];
            const float jji = S[(size_t)n * L + r];
            const double v = (double)(float)(jij + jji) * 0.5;  // plmDCA.hpp:690
            p = put_fixed8(p, v);
            *p++ = ' ';
            p = put_uint(p, ri);
            *p++ = ' ';
            p = put_uint(p, (uint64_t)((int64_t)translation[n] + base));
            *p++ = '\n';
            if (p >= end) {
                if (fwrite(buf.data(), 1, p - buf.data(), fp) != (size_t)(p - buf.data())) rc = SDCA_E_IO;
                p = buf.data();
            }
        }
    }
    if (rc == SDCA_OK && p > buf.data() && fwrite(buf.data(), 1, p - buf.data(), fp) != (size_t)(p - buf.data()))
        rc = SDCA_E_IO;
    if (fclose(fp) != 0) rc = SDCA_E_IO;
    if (rc != SDCA_OK) sdca_fail(rc, "writing \"%s\" failed", path);
    return rc;
}

// Two-alignment scan output: row k of S against every locus of the second alignment,
// the float score as printed by the reference's fixed 8-digit stream.
int sdca_write_scan_couplings(const char *path, const float *S, int32_t K, int32_t L2, const int32_t *translation1,
                              const int32_t *translation2, int32_t base)
{
    if (!path || !S || !translation1 || !translation2 || K < 0 || L2 < 0)
        return sdca_fail(SDCA_E_ARG, "sdca_write_scan_couplings: null argument or negative size");
    FILE *fp = fopen(path, "wb");
    if (!fp) return sdca_fail(SDCA_E_IO, "cannot open \"%s\" for writing", path);
    std::vector<char> buf(1 << 22);
    char *p = buf.data();
    char *const end = buf.data() + buf.size() - 128;
    int rc = SDCA_OK;
    for (int k = 0; k < K && rc == SDCA_OK; ++k) {
        const uint64_t ri = (uint64_t)((int64_t)translation1[k] + base);
        for (int n = 0; n < L2; ++n) {
            p = put_fixed8(p, (double)S[(size_t)k * L2 + n]);
            *p++ = ' ';
            p = put_uint(p, ri);
            *p++ = ' ';
            p = put_uint(p, (uint64_t)((int64_t)translation2[n] + base));
            *p++ = '\n';
            if (p >= end) {
                if (fwrite(buf.data(), 1, p - buf.data(), fp) != (size_t)(p - buf.data())) rc = SDCA_E_IO;
                p = buf.data();
            }
        }
    }
    if (rc == SDCA_OK && p > buf.data() && fwrite(buf.data(), 1, p - buf.data(), fp) != (size_t)(p - buf.data()))
        rc = SDCA_E_IO;
    if (fclose(fp) != 0) rc = SDCA_E_IO;
    if (rc != SDCA_OK) sdca_fail(rc, "writing \"%s\" failed", path);
    return rc;
}

// The highest symmetrised scores only: the pairs of sdca_write_couplings with a value of
// at least min_score, the top_n largest of them (all if top_n == 0), largest first, equal
// values in the order of the full file.  One pass over the pairs with a bounded min-heap,
// so a genome-scale matrix needs no sort of its L(L-1)/2 values.
int sdca_write_couplings_top(const char *path, const float *S, int32_t L, const int32_t *translation, int32_t base,
                             int64_t top_n, double min_score)
{
    if (!path || !S || !translation || L < 0 || top_n < 0 || std::isnan(min_score))
        return sdca_fail(SDCA_E_ARG, "sdca_write_couplings_top: null argument, L < 0, top_n < 0 or NaN min_score");
    struct Pair {
        double v;
        int64_t ord;  // position in the full file
        int32_t r, n;
    };
    // a before b in the output
    auto first = [](const Pair &a, const Pair &b) { return a.v > b.v || (a.v == b.v && a.ord < b.ord); };
    std::vector<Pair> keep;
    int64_t ord = 0;
    for (int r = 0; r < L; ++r)
        for (int n = 0; n < r; ++n, ++ord) {
            const double v = (double)(float)(S[(size_t)r * L + n] + S[(size_t)n * L + r]) * 0.5;
            if (!(v >= min_score)) continue;  // also drops NaN
            const Pair x{v, ord, r, n};
            if (top_n == 0 || (int64_t)keep.size() < top_n) {
                keep.push_back(x);
                if (top_n > 0) std::push_heap(keep.begin(), keep.end(), first);  // front: the last in order
            } else if (first(x, keep.front())) {
                std::pop_heap(keep.begin(), keep.end(), first);
                keep.back() = x;
                std::push_heap(keep.begin(), keep.end(), first);
            }
        }
    std::sort(keep.begin(), keep.end(), first);
    FILE *fp = fopen(path, "wb");
    if (!fp) return sdca_fail(SDCA_E_IO, "cannot open \"%s\" for writing", path);
    int rc = SDCA_OK;
    char line[128];  // a float score prints in at most 50 characters, each index in 20
    for (const Pair &x : keep) {
        char *p = put_fixed8(line, x.v);
        *p++ = ' ';
        p = put_uint(p, (uint64_t)((int64_t)translation[x.r] + base));
        *p++ = ' ';
        p = put_uint(p, (uint64_t)((int64_t)translation[x.n] + base));
        *p++ = '\n';
        if (fwrite(line, 1, p - line, fp) != (size_t)(p - line)) {
            rc = SDCA_E_IO;
            break;
        }
    }
    if (fclose(fp) != 0) rc = SDCA_E_IO;
    if (rc != SDCA_OK) sdca_fail(rc, "writing \"%s\" failed", path);
    return rc;
}

}  // extern "C"
