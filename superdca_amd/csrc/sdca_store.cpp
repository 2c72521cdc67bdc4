// sdca_store.cpp -- file-backed score and coupling-block stores and their streaming writers.
//
// SuperDCA keeps every score (or, with --norm-of-mean-scoring, every q x q coupling block)
// of the run in memory and writes the output from there (SuperDCA/include/plmDCA.hpp:
// 86-190 storage, 589-697 writer).  At genome scale that is 35 GB of scores, or 565 GB and
// more of blocks, so here the rows go to a file as the solve produces them and the writers
// stream the file in tiles of rows: host memory stays at a few tiles of rows.
//
// A store holds `rows` x `cols` entries of `entry` floats, row-major, in an unlinked file
// (the space is returned when the store is destroyed or the process ends).
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sdca.h"
#include "sdca_format.hpp"

// The norm-of-mean arithmetic below restates float expressions of the reference in their
// order; no contraction into fused multiply-adds.
#pragma clang fp contract(off)

int sdca_fail(int code, const char *fmt, ...);  // sdca_api.cpp

namespace sdca {
size_t ref_index(int L, int q, int j, int a, int s);  // sdca_host.cpp
}

struct sdca_store {
    int fd = -1;
    int64_t rows = 0, cols = 0;
    int32_t entry = 1;
    ~sdca_store()
    {
        if (fd >= 0) close(fd);
    }
};

namespace {

bool pwrite_all(int fd, const void *buf, size_t n, int64_t off)
{
    const char *p = static_cast<const char *>(buf);
    while (n > 0) {
        const ssize_t w = pwrite(fd, p, n, (off_t)off);
        if (w < 0 && errno == EINTR) continue;
        if (w <= 0) return false;
        p += w;
        n -= (size_t)w;
        off += w;
    }
    return true;
}

bool pread_all(int fd, void *buf, size_t n, int64_t off)
{
    char *p = static_cast<char *>(buf);
    while (n > 0) {
        const ssize_t r = pread(fd, p, n, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        p += r;
        n -= (size_t)r;
        off += r;
    }
    return true;
}

// Entries [c0, c1) of row r into out.
bool read_entries(const sdca_store *s, int64_t r, int64_t c0, int64_t c1, float *out)
{
    if (c1 <= c0) return true;
    const int64_t e = s->entry;
    return pread_all(s->fd, out, (size_t)((c1 - c0) * e) * sizeof(float), ((r * s->cols + c0) * e) * (int64_t)sizeof(float));
}

// Rows of a tile: as many as keep the row band and the column band within ~256 MB each.
int64_t tile_rows(const sdca_store *s, int32_t tile)
{
    if (tile > 0) return tile;
    const int64_t per_row = std::max<int64_t>(1, s->cols * s->entry * (int64_t)sizeof(float));
    return std::max<int64_t>(1, std::min<int64_t>(s->rows, (int64_t(256) << 20) / per_row));
}

// Buffered line output.
struct Out {
    FILE *fp = nullptr;
    std::vector<char> buf = std::vector<char>(1 << 22);
    char *p = nullptr;
    bool ok = true;
    explicit Out(FILE *f) : fp(f), p(buf.data()) {}
    char *reserve()
    {
        if (p > buf.data() + buf.size() - 256) flush();
        return p;
    }
    void flush()
    {
        const size_t n = (size_t)(p - buf.data());
        if (n && fwrite(buf.data(), 1, n, fp) != n) ok = false;
        p = buf.data();
    }
};

char *put_pair(char *p, double v, uint64_t ri, uint64_t ni)
{
    p = sdca::put_fixed8(p, v);
    *p++ = ' ';
    p = sdca::put_uint(p, ri);
    *p++ = ' ';
    p = sdca::put_uint(p, ni);
    return p;
}

// The largest values in the order of sdca_write_couplings_top: value descending, then the
// position in the full file.
struct TopKeeper {
    struct Pair {
        double v;
        int64_t ord;
        int32_t r, n;
    };
    int64_t top_n;
    double min_score;
    std::vector<Pair> keep;
    static bool first(const Pair &a, const Pair &b) { return a.v > b.v || (a.v == b.v && a.ord < b.ord); }
    void add(double v, int64_t ord, int32_t r, int32_t n)
    {
        if (!(v >= min_score)) return;  // also drops NaN
        const Pair x{v, ord, r, n};
        if (top_n == 0 || (int64_t)keep.size() < top_n) {
            keep.push_back(x);
            if (top_n > 0) std::push_heap(keep.begin(), keep.end(), first);
        } else if (first(x, keep.front())) {
            std::pop_heap(keep.begin(), keep.end(), first);
            keep.back() = x;
            std::push_heap(keep.begin(), keep.end(), first);
        }
    }
};

// Ising gauge of a float block in the reference's float arithmetic and order
// (plmDCA_utility.hpp:31-54 on a float matrix view; mean, repmat and the element
// operators of Matrix_math.hpp:51-76,215-374): row means r[i], column means c[j], the mean
// m of the row means, G[i][j] = ((M[i][j] - r[j]) - c[i]) + m.
void gauge_float(const float *M, int q, float *G)
{
    float rm[8], cm[8];
    for (int i = 0; i < q; ++i) {
        float a = 0.f;
        for (int j = 0; j < q; ++j) a = a + M[i * q + j];
        rm[i] = a / (float)q;
    }
    for (int j = 0; j < q; ++j) {
        float a = 0.f;
        for (int i = 0; i < q; ++i) a = a + M[i * q + j];
        cm[j] = a / (float)q;
    }
    float mm = 0.f;
    for (int i = 0; i < q; ++i) mm += rm[i];
    mm = mm / (float)q;
    for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j) G[i * q + j] = ((M[i * q + j] - rm[j]) - cm[i]) + mm;
}

// frobenius_norm(a, exclude) of a float matrix (Matrix_math.hpp:132-178): the float sum
// takes std::pow(a, 2), a double, in each step; the square root is taken in float.
float frobenius_float(const float *G, int q, int exclude)
{
    float sum = 0.f;
    for (int i = 0; i < q; ++i) {
        if (i == exclude) continue;
        for (int j = 0; j < q; ++j) {
            if (j == exclude) continue;
            const double a = (double)G[i * q + j];
            sum = (float)((double)sum + a * a);
        }
    }
    return std::sqrt(sum);
}

bool valid_gap(int q, int gap) { return gap >= -1 && gap < q; }

}  // namespace

extern "C" {

int sdca_store_create(sdca_store **out, const char *path, int64_t rows, int64_t cols, int32_t entry)
{
    if (!out) return sdca_fail(SDCA_E_ARG, "out is null");
    *out = nullptr;
    if (!path || rows < 1 || cols < 1 || entry < 1 || entry > 64)
        return sdca_fail(SDCA_E_ARG, "sdca_store_create: need a path, rows, cols >= 1 and entry in [1, 64]");
    const int fd = open(path, O_RDWR | O_CREAT | O_EXCL, 0600);
    if (fd < 0) return sdca_fail(SDCA_E_IO, "cannot create \"%s\": %s", path, strerror(errno));
    unlink(path);  // the space lives as long as the store
    const int64_t bytes = rows * cols * entry * (int64_t)sizeof(float);
    if (ftruncate(fd, (off_t)bytes) != 0) {
        close(fd);
        return sdca_fail(SDCA_E_IO, "cannot size the store \"%s\" to %lld bytes", path, (long long)bytes);
    }
    auto *s = new sdca_store();
    s->fd = fd;
    s->rows = rows;
    s->cols = cols;
    s->entry = entry;
    *out = s;
    return SDCA_OK;
}

void sdca_store_destroy(sdca_store *s) { delete s; }

int sdca_store_put_rows(sdca_store *s, int64_t r0, int64_t n, const float *rows)
{
    if (!s || (n > 0 && !rows) || r0 < 0 || n < 0 || r0 + n > s->rows)
        return sdca_fail(SDCA_E_ARG, "sdca_store_put_rows: rows [%lld, %lld) outside the store", (long long)r0,
                         (long long)(r0 + n));
    const size_t bytes = (size_t)(n * s->cols * s->entry) * sizeof(float);
    if (!pwrite_all(s->fd, rows, bytes, r0 * s->cols * s->entry * (int64_t)sizeof(float)))
        return sdca_fail(SDCA_E_IO, "writing store rows failed: %s", strerror(errno));
    return SDCA_OK;
}

int sdca_store_put_blocks(sdca_store *s, int64_t row, const double *x, int32_t L, int32_t q, int32_t r,
                          const int32_t *col_loci)
{
    if (!s || !x || !col_loci || row < 0 || row >= s->rows || q < 2 || q > 8 || s->entry != q * q || r < 0 ||
        r >= L)
        return sdca_fail(SDCA_E_ARG, "sdca_store_put_blocks: bad arguments (the store entry must be q*q)");
    const int qq = q * q;
    std::vector<float> out((size_t)s->cols * qq, 0.f);
    for (int64_t u = 0; u < s->cols; ++u) {
        const int n = col_loci[u];
        if (n < 0 || n >= L) return sdca_fail(SDCA_E_ARG, "sdca_store_put_blocks: locus %d outside [0, %d)", n, L);
        if (n == r) continue;
        float *M = out.data() + u * qq;
        // plmDCA.hpp:339-361: the block J_r[n](a, s) (a: state at n, s: state at r) as float,
        // transposed for n < r, so that a pair's two blocks both run [state at the larger
        // locus][state at the smaller]
        for (int a = 0; a < q; ++a) {
            const size_t o = sdca::ref_index(L, q, n, a, 0);
            for (int t = 0; t < q; ++t) {
                const float v = (float)x[o + t];
                if (n < r) M[t * q + a] = v;
                else M[a * q + t] = v;
            }
        }
    }
    if (!pwrite_all(s->fd, out.data(), out.size() * sizeof(float), row * s->cols * qq * (int64_t)sizeof(float)))
        return sdca_fail(SDCA_E_IO, "writing store blocks failed: %s", strerror(errno));
    return SDCA_OK;
}

int sdca_store_write_couplings(const sdca_store *s, const char *path_all, const char *path_top,
                               const int32_t *translation, int32_t base, int64_t top_n, double min_score, int32_t tile)
{
    if (!s || !translation || s->entry != 1 || s->rows != s->cols || top_n < 0 || std::isnan(min_score) || tile < 0)
        return sdca_fail(SDCA_E_ARG, "sdca_store_write_couplings: needs a square store of scores");
    if (!path_all && !path_top) return SDCA_OK;
    FILE *fa = path_all ? fopen(path_all, "wb") : nullptr;
    if (path_all && !fa) return sdca_fail(SDCA_E_IO, "cannot open \"%s\" for writing", path_all);
    Out out(fa);
    TopKeeper top{top_n, min_score, {}};
    const int64_t K = s->rows, T = tile_rows(s, tile);
    std::vector<float> rowband, colband;
    int rc = SDCA_OK;
    int64_t ord = 0;
    for (int64_t r0 = 0; r0 < K && rc == SDCA_OK; r0 += T) {
        const int64_t r1 = std::min(K, r0 + T), nt = r1 - r0;
        rowband.assign((size_t)(nt * r1), 0.f);  // S[r][0, r1) of the tile's rows
        colband.assign((size_t)(r1 * nt), 0.f);  // S[n][r0, r1) for n < r1
        for (int64_t r = r0; r < r1 && rc == SDCA_OK; ++r)
            if (!read_entries(s, r, 0, r1, rowband.data() + (r - r0) * r1)) rc = SDCA_E_IO;
        for (int64_t n = 0; n < r1 && rc == SDCA_OK; ++n)
            if (!read_entries(s, n, r0, r1, colband.data() + n * nt)) rc = SDCA_E_IO;
        for (int64_t r = r0; r < r1 && rc == SDCA_OK; ++r) {
            const uint64_t ri = (uint64_t)((int64_t)translation[r] + base);
            for (int64_t n = 0; n < r; ++n, ++ord) {
                const float jij = rowband[(r - r0) * r1 + n];
                const float jji = colband[n * nt + (r - r0)];
                const double v = (double)(float)(jij + jji) * 0.5;  // plmDCA.hpp:690
                if (fa) {
                    char *p = put_pair(out.reserve(), v, ri, (uint64_t)((int64_t)translation[n] + base));
                    *p++ = '\n';
                    out.p = p;
                }
                if (path_top) top.add(v, ord, (int32_t)r, (int32_t)n);
            }
        }
    }
    if (fa) {
        out.flush();
        if (!out.ok) rc = SDCA_E_IO;
        if (fclose(fa) != 0) rc = SDCA_E_IO;
    }
    if (rc != SDCA_OK) return sdca_fail(rc, "streaming the couplings failed (store read or output write)");
    if (path_top) {
        std::sort(top.keep.begin(), top.keep.end(), TopKeeper::first);
        FILE *ft = fopen(path_top, "wb");
        if (!ft) return sdca_fail(SDCA_E_IO, "cannot open \"%s\" for writing", path_top);
        Out o2(ft);
        for (const auto &x : top.keep) {
            char *p = put_pair(o2.reserve(), x.v, (uint64_t)((int64_t)translation[x.r] + base),
                               (uint64_t)((int64_t)translation[x.n] + base));
            *p++ = '\n';
            o2.p = p;
        }
        o2.flush();
        if (!o2.ok) rc = SDCA_E_IO;
        if (fclose(ft) != 0) rc = SDCA_E_IO;
        if (rc != SDCA_OK) return sdca_fail(rc, "writing \"%s\" failed", path_top);
    }
    return SDCA_OK;
}

int sdca_store_write_norm_of_mean(const sdca_store *s, const char *path, const int32_t *translation, int32_t base,
                                  int32_t q, int32_t gap, int32_t tile)
{
    if (!s || !path || !translation || q < 2 || q > 8 || s->entry != q * q || s->rows != s->cols || !valid_gap(q, gap) ||
        tile < 0)
        return sdca_fail(SDCA_E_ARG, "sdca_store_write_norm_of_mean: needs a square store of q*q blocks");
    FILE *fp = fopen(path, "wb");
    if (!fp) return sdca_fail(SDCA_E_IO, "cannot open \"%s\" for writing", path);
    Out out(fp);
    const int qq = q * q;
    const int64_t K = s->rows, T = tile_rows(s, tile);
    std::vector<float> rowband, colband;
    int rc = SDCA_OK;
    for (int64_t r0 = 0; r0 < K && rc == SDCA_OK; r0 += T) {
        const int64_t r1 = std::min(K, r0 + T), nt = r1 - r0;
        rowband.assign((size_t)(nt * r1 * qq), 0.f);
        colband.assign((size_t)(r1 * nt * qq), 0.f);
        for (int64_t r = r0; r < r1 && rc == SDCA_OK; ++r)
            if (!read_entries(s, r, 0, r1, rowband.data() + (r - r0) * r1 * qq)) rc = SDCA_E_IO;
        for (int64_t n = 0; n < r1 && rc == SDCA_OK; ++n)
            if (!read_entries(s, n, r0, r1, colband.data() + n * nt * qq)) rc = SDCA_E_IO;
        for (int64_t r = r0; r < r1 && rc == SDCA_OK; ++r) {
            const uint64_t ri = (uint64_t)((int64_t)translation[r] + base);
            for (int64_t n = 0; n < r; ++n) {
                // plmDCA.hpp:640-664: the gauged blocks of (r, n) and (n, r), their norms and
                // the norm of their mean, all in float
                float Gij[64], Gji[64], Gm[64];
                gauge_float(rowband.data() + ((r - r0) * r1 + n) * qq, q, Gij);
                gauge_float(colband.data() + (n * nt + (r - r0)) * qq, q, Gji);
                for (int e = 0; e < qq; ++e) Gm[e] = (Gij[e] + Gji[e]) * 0.5f;
                const float nij = frobenius_float(Gij, q, gap), nji = frobenius_float(Gji, q, gap);
                const float nm = frobenius_float(Gm, q, gap);
                char *p = put_pair(out.reserve(), (double)nm, ri, (uint64_t)((int64_t)translation[n] + base));
                *p++ = ' ';
                p = sdca::put_fixed8(p, (double)nij);
                *p++ = ' ';
                p = sdca::put_fixed8(p, (double)nji);
                *p++ = ' ';
                p = sdca::put_fixed8(p, (double)(nij + nji) * 0.5);
                *p++ = '\n';
                out.p = p;
            }
        }
    }
    out.flush();
    if (!out.ok) rc = SDCA_E_IO;
    if (fclose(fp) != 0) rc = SDCA_E_IO;
    if (rc != SDCA_OK) return sdca_fail(rc, "writing \"%s\" failed", path);
    return SDCA_OK;
}

int sdca_store_write_norm_of_mean_scan(const sdca_store *s, const char *path, const int32_t *translation1,
                                       const int32_t *translation2, int32_t base, int32_t q, int32_t gap)
{
    if (!s || !path || !translation1 || !translation2 || q < 2 || q > 8 || s->entry != q * q || !valid_gap(q, gap))
        return sdca_fail(SDCA_E_ARG, "sdca_store_write_norm_of_mean_scan: needs a store of q*q blocks");
    FILE *fp = fopen(path, "wb");
    if (!fp) return sdca_fail(SDCA_E_IO, "cannot open \"%s\" for writing", path);
    Out out(fp);
    const int qq = q * q;
    std::vector<float> row((size_t)s->cols * qq);
    int rc = SDCA_OK;
    for (int64_t r = 0; r < s->rows && rc == SDCA_OK; ++r) {
        if (!read_entries(s, r, 0, s->cols, row.data())) {
            rc = SDCA_E_IO;
            break;
        }
        const uint64_t ri = (uint64_t)((int64_t)translation1[r] + base);
        for (int64_t n = 0; n < s->cols; ++n) {
            // plmDCA.hpp:625-638: the gauged block's norm, not symmetrised
            float G[64];
            gauge_float(row.data() + n * qq, q, G);
            char *p = put_pair(out.reserve(), (double)frobenius_float(G, q, gap), ri,
                               (uint64_t)((int64_t)translation2[n] + base));
            *p++ = '\n';
            out.p = p;
        }
    }
    out.flush();
    if (!out.ok) rc = SDCA_E_IO;
    if (fclose(fp) != 0) rc = SDCA_E_IO;
    if (rc != SDCA_OK) return sdca_fail(rc, "writing \"%s\" failed", path);
    return SDCA_OK;
}

}  // extern "C"
