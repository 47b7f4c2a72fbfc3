// qc_host.hip -- run-time quasi-cyclic codes: construction from a base
// matrix, alist reading and writing, quasi-cyclic structure detection (host
// code; the decoder kernel is generic.hip).
//
// The alist format is the one the reference writes and reads
// (python_ldpc/ldpc_utils.py:157-240, python_ldpc/ldpc_decoder.py:20-59):
//   N M / max column weight, max row weight / N column weights / M row
//   weights / N lines of 1-based check indices per variable, 0-padded /
//   M lines of 1-based variable indices per check, 0-padded
// with every number followed by one space.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "errors.hpp"
#include "qc_host.hpp"

namespace {

int qfail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    wl::set_error_text(buf);
    return code;
}

int gcd_i(int a, int b) { return b == 0 ? a : gcd_i(b, a % b); }

// rows [M] of ascending 0-based variable indices -> base matrix at lifting z,
// or false if the rows are not z x z circulant permutation blocks in
// layer-major order
bool detect_qc(const std::vector<std::vector<int>>& rows, int N, int M, int z, std::vector<int16_t>& base)
{
    if (z < 1 || N % z || M % z) return false;
    const int nb = N / z, mb = M / z;
    base.assign((size_t)mb * nb, (int16_t)-1);
    std::vector<int> expect;
    for (int l = 0; l < mb; l++) {
        const std::vector<int>& r0 = rows[(size_t)l * z];
        for (int c : r0) {
            const int b = c / z;
            if (base[(size_t)l * nb + b] != -1) return false;   // two entries of one block in a row
            base[(size_t)l * nb + b] = (int16_t)(c % z);
        }
        for (int r = 1; r < z; r++) {
            expect.clear();
            for (int b = 0; b < nb; b++) {
                const int s = base[(size_t)l * nb + b];
                if (s >= 0) expect.push_back(b * z + (r + s) % z);
            }
            if (expect != rows[(size_t)l * z + r]) return false;
        }
    }
    return true;
}

}  // namespace

namespace wl {

int qc_build(int nb, int mb, int z, const int16_t* base, wl_qc_code& q)
{
    if (nb < 2 || mb < 1 || mb >= nb) return qfail(WL_EINVAL, "base matrix %d x %d: need 1 <= mb < nb", mb, nb);
    if (mb > kGenMaxLayers) return qfail(WL_ENOTSUP, "%d base rows: at most %d supported", mb, kGenMaxLayers);
    if (z < 1 || z > kGenMaxZ) return qfail(WL_ENOTSUP, "lifting size %d: 1..%d supported", z, kGenMaxZ);
    if ((long)nb * z > 32767) return qfail(WL_ENOTSUP, "code length %d too large", nb * z);
    if (!base) return qfail(WL_EINVAL, "null base matrix");
    q.nb = nb;
    q.mb = mb;
    q.z = z;
    q.base.assign(base, base + (size_t)mb * nb);
    int c = 0, edges = 0;
    for (int l = 0; l < mb; l++) {
        int dg = 0;
        for (int b = 0; b < nb; b++) {
            const int s = base[(size_t)l * nb + b];
            if (s < -1 || s >= z) return qfail(WL_EINVAL, "shift %d at (%d, %d) outside -1..%d", s, l, b, z - 1);
            dg += s >= 0;
        }
        if (dg == 0) return qfail(WL_EINVAL, "base row %d is empty", l);
        if (dg > kGenMaxDeg) return qfail(WL_ENOTSUP, "base row %d has %d blocks: at most %d", l, dg, kGenMaxDeg);
        c = std::max(c, dg);
        edges += dg * z;
    }
    if (gen_lds_bytes(nb * z, edges) > (size_t)kGenMaxLds)
        return qfail(WL_ENOTSUP, "code needs %zu bytes of LDS per frame (limit %d)", gen_lds_bytes(nb * z, edges),
                     kGenMaxLds);
    wl_code_info& ci = q.info;
    ci.rate = -1;
    ci.z = z;
    ci.n = nb * z;
    ci.m = mb * z;
    ci.k = ci.n - ci.m;
    ci.c = c;
    ci.mb = mb;
    ci.d = -1;
    ci.x = -1;
    q.table.assign((size_t)ci.m * c, (int16_t)-1);
    GenDesc& d = q.desc;
    memset(&d, 0, sizeof d);
    d.z = z;
    d.nb = nb;
    d.mb = mb;
    d.n = ci.n;
    d.m = ci.m;
    d.edges = edges;
    int off = 0;
    for (int l = 0; l < mb; l++) {
        int j = 0;
        for (int b = 0; b < nb; b++) {
            const int s = base[(size_t)l * nb + b];
            if (s < 0) continue;
            d.blk[l][j] = (int16_t)b;
            d.sh[l][j] = (int16_t)s;
            for (int r = 0; r < z; r++) q.table[((size_t)l * z + r) * c + j] = (int16_t)(b * z + (r + s) % z);
            j++;
        }
        d.deg[l] = j;
        d.moff[l] = off;
        off += j * z;
    }
    // dual-diagonal parity part (802.16e / 802.11n): column kb holds shifts
    // p, 0, p in base rows 0, x, mb-1 and nothing else; column kb+1+i holds
    // shift 0 in rows i and i+1 only
    const int kb = nb - mb;
    bool dual = mb >= 3;
    int x = -1;
    for (int l = 0; dual && l < mb; l++) {
        const int s = base[(size_t)l * nb + kb];
        if (l == 0 || l == mb - 1) {
            dual = s > 0 && s == base[(size_t)(mb - 1) * nb + kb];
        } else if (s == 0) {
            if (x >= 0) dual = false;
            x = l;
        } else if (s != -1) {
            dual = false;
        }
    }
    for (int i = 0; dual && i + 1 < mb; i++)
        for (int l = 0; l < mb; l++) {
            const int s = base[(size_t)l * nb + kb + 1 + i];
            if ((l == i || l == i + 1) ? s != 0 : s != -1) dual = false;
        }
    if (dual && x > 0) {
        ci.x = x;
        ci.d = base[kb];
    }
    return WL_OK;
}

}  // namespace wl

extern "C" {

int wl_qc_create(int nb, int mb, int z, const int16_t* base, wl_qc_code** out)
{
    if (!out) return qfail(WL_EINVAL, "null output");
    *out = nullptr;
    wl_qc_code* q = new (std::nothrow) wl_qc_code();
    if (!q) return qfail(WL_ENOMEM, "out of host memory");
    const int rc = wl::qc_build(nb, mb, z, base, *q);
    if (rc) {
        delete q;
        return rc;
    }
    *out = q;
    return WL_OK;
}

int wl_qc_from_alist(const char* text, size_t len, int z, wl_qc_code** out)
{
    if (!out || !text) return qfail(WL_EINVAL, "null argument");
    *out = nullptr;
    std::vector<long> tok;
    {
        std::string s(text, len);
        const char* p = s.c_str();
        char* end = nullptr;
        for (;;) {
            while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
            if (!*p) break;
            const long v = strtol(p, &end, 10);
            if (end == p) return qfail(WL_EINVAL, "alist: unexpected character '%c'", *p);
            tok.push_back(v);
            p = end;
        }
    }
    size_t t = 0;
    auto next = [&](long& v) -> bool {
        if (t >= tok.size()) return false;
        v = tok[t++];
        return true;
    };
    long N, M, wc, wr;
    // a column has at most M entries and a row at most N, so with N <= 32767
    // every size below stays far from overflow
    if (!next(N) || !next(M) || !next(wc) || !next(wr) || N < 2 || M < 1 || M >= N || N > 32767 || wc < 1 ||
        wr < 1 || wc > M || wr > N)
        return qfail(WL_EINVAL, "alist: bad header");
    const size_t need = 4 + (size_t)N + (size_t)M + (size_t)N * wc + (size_t)M * wr;
    if (tok.size() < need) return qfail(WL_EINVAL, "alist: %zu numbers, header needs %zu", tok.size(), need);
    t += (size_t)N + (size_t)M;   // weights are implied by the lists
    t += (size_t)N * wc;          // column lists: the row lists carry the same graph
    std::vector<std::vector<int>> rows((size_t)M);
    for (long m = 0; m < M; m++) {
        for (long j = 0; j < wr; j++) {
            long v = 0;
            if (!next(v)) return qfail(WL_EINVAL, "alist: check list %ld ends early", m + 1);
            if (v < 0 || v > N) return qfail(WL_EINVAL, "alist: variable %ld out of range in check %ld", v, m + 1);
            if (v) rows[(size_t)m].push_back((int)v - 1);
        }
        std::sort(rows[(size_t)m].begin(), rows[(size_t)m].end());
        if (std::adjacent_find(rows[(size_t)m].begin(), rows[(size_t)m].end()) != rows[(size_t)m].end())
            return qfail(WL_EINVAL, "alist: check %ld lists a variable twice", m + 1);
    }
    std::vector<int16_t> base;
    int zz = 0;
    if (z > 0) {
        if (!detect_qc(rows, (int)N, (int)M, z, base))
            return qfail(WL_EINVAL, "alist: not quasi-cyclic with lifting size %d", z);
        zz = z;
    } else {
        const int g = gcd_i((int)N, (int)M);
        for (int c = g; c >= 1; c--)
            if (g % c == 0 && c <= wl::kGenMaxZ && detect_qc(rows, (int)N, (int)M, c, base)) {
                zz = c;
                break;
            }
        if (!zz) return qfail(WL_EINVAL, "alist: no quasi-cyclic structure found");
    }
    return wl_qc_create((int)N / zz, (int)M / zz, zz, base.data(), out);
}

int wl_qc_destroy(wl_qc_code* q)
{
    delete q;
    return WL_OK;
}

int wl_qc_params(const wl_qc_code* q, wl_code_info* out)
{
    if (!q || !out) return qfail(WL_EINVAL, "null argument");
    *out = q->info;
    return WL_OK;
}

int wl_qc_base(const wl_qc_code* q, int* nb, int* mb, int16_t* out, size_t out_len)
{
    if (!q) return qfail(WL_EINVAL, "null code");
    if (nb) *nb = q->nb;
    if (mb) *mb = q->mb;
    if (!out) return WL_OK;
    if (out_len < q->base.size()) return qfail(WL_EINVAL, "base matrix needs %zu entries", q->base.size());
    memcpy(out, q->base.data(), q->base.size() * sizeof(int16_t));
    return WL_OK;
}

int wl_qc_table(const wl_qc_code* q, int16_t* out, size_t out_len)
{
    if (!q || !out) return qfail(WL_EINVAL, "null argument");
    if (out_len < q->table.size()) return qfail(WL_EINVAL, "table needs %zu entries", q->table.size());
    memcpy(out, q->table.data(), q->table.size() * sizeof(int16_t));
    return WL_OK;
}

int wl_qc_alist(const wl_qc_code* q, char* out, size_t cap, size_t* len)
{
    if (!q || !len) return qfail(WL_EINVAL, "null argument");
    const int N = q->info.n, M = q->info.m, c = q->info.c;
    std::vector<std::vector<int>> cols((size_t)N);
    int wr = 0;
    for (int m = 0; m < M; m++) {
        int w = 0;
        for (int j = 0; j < c; j++) {
            const int v = q->table[(size_t)m * c + j];
            if (v < 0) continue;
            cols[(size_t)v].push_back(m);
            w++;
        }
        wr = std::max(wr, w);
    }
    int wc = 0;
    for (auto& col : cols) wc = std::max(wc, (int)col.size());
    std::string s;
    char buf[32];
    auto put = [&](int v) {
        snprintf(buf, sizeof buf, "%d ", v);
        s += buf;
    };
    snprintf(buf, sizeof buf, "%d %d\n", N, M);
    s += buf;
    snprintf(buf, sizeof buf, "%d %d\n", wc, wr);
    s += buf;
    for (auto& col : cols) put((int)col.size());
    s += "\n";
    for (int m = 0; m < M; m++) {
        int w = 0;
        for (int j = 0; j < c; j++) w += q->table[(size_t)m * c + j] >= 0;
        put(w);
    }
    s += "\n";
    for (auto& col : cols) {
        for (int j = 0; j < wc; j++) put(j < (int)col.size() ? col[(size_t)j] + 1 : 0);
        s += "\n";
    }
    for (int m = 0; m < M; m++) {
        std::vector<int> row;
        for (int j = 0; j < c; j++)
            if (q->table[(size_t)m * c + j] >= 0) row.push_back(q->table[(size_t)m * c + j]);
        std::sort(row.begin(), row.end());
        for (int j = 0; j < wr; j++) put(j < (int)row.size() ? row[(size_t)j] + 1 : 0);
        s += "\n";
    }
    *len = s.size();
    if (!out) return WL_OK;
    if (cap < s.size() + 1) return qfail(WL_EINVAL, "alist needs %zu bytes", s.size() + 1);
    memcpy(out, s.c_str(), s.size() + 1);
    return WL_OK;
}

}  // extern "C"
