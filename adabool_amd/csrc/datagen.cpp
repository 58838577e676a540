// datagen.cpp -- seeded synthetic filters and queries, text loaders, stats.
//
// The generators use their own PRNG (splitmix64 seeding xoshiro256**) so the
// same seed gives the same data on every host and thread count: every row and
// every query draws from a stream derived from (seed, row or query index).
// Zipf(s) popularity over nkeys ranks is sampled with a Vose alias table; ranks
// map to key ids through a permutation seeded by perm_seed, shared by the
// filter and the queries built with the same perm_seed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <unordered_set>

#include <dirent.h>
#include <sys/stat.h>

#include "internal.h"

namespace ab {

static inline uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

static uint64_t stream_seed(uint64_t seed, uint64_t idx, uint64_t salt)
{
    uint64_t x = seed ^ (salt * 0xD1B54A32D192ED03ull);
    uint64_t a = splitmix64(x);
    x = a ^ (idx * 0x9E3779B97F4A7C15ull);
    return splitmix64(x);
}

struct ZipfAlias {
    std::vector<double> prob;
    std::vector<uint32_t> alias;
    std::vector<uint32_t> perm;  // rank -> key
    ZipfAlias(uint32_t n, double s, uint64_t perm_seed)
    {
        prob.resize(n);
        alias.resize(n);
        std::vector<double> w(n);
        double sum = 0;
        for (uint32_t r = 0; r < n; ++r) {
            w[r] = std::pow((double)r + 1.0, -s);
            sum += w[r];
        }
        std::vector<uint32_t> small, large;
        small.reserve(n);
        large.reserve(n);
        for (uint32_t r = 0; r < n; ++r) {
            w[r] = w[r] * n / sum;
            (w[r] < 1.0 ? small : large).push_back(r);
        }
        while (!small.empty() && !large.empty()) {
            uint32_t a = small.back(), b = large.back();
            small.pop_back();
            prob[a] = w[a];
            alias[a] = b;
            w[b] = (w[b] + w[a]) - 1.0;
            if (w[b] < 1.0) {
                large.pop_back();
                small.push_back(b);
            }
        }
        for (uint32_t r : large) prob[r] = 1.0, alias[r] = r;
        for (uint32_t r : small) prob[r] = 1.0, alias[r] = r;
        perm.resize(n);
        for (uint32_t r = 0; r < n; ++r) perm[r] = r;
        Rng g(stream_seed(perm_seed, 0, 0x5EED));
        for (uint32_t r = n; r > 1; --r) std::swap(perm[r - 1], perm[g.below(r)]);
    }
    uint32_t sample(Rng &g) const
    {
        uint32_t i = g.below((uint32_t)prob.size());
        uint32_t r = g.uniform() < prob[i] ? i : alias[i];
        return perm[r];
    }
};

static uint32_t draw_len(Rng &g, int dist, double p1, double p2, uint32_t lo, uint32_t hi)
{
    double x;
    if (dist == 0) return lo + g.below(hi - lo + 1);
    if (dist == 1) {
        x = std::round(std::exp(std::log(p1) + p2 * g.normal()));
    } else {
        // continuous Pareto on [lo, hi+1) with density ~ x^-p1, floored
        double a = 1.0 - p1, u = g.uniform();
        double l = std::pow((double)lo, a), h = std::pow((double)hi + 1.0, a);
        x = std::floor(std::pow(l + u * (h - l), 1.0 / a));
    }
    if (!(x >= lo)) x = lo;
    if (x > hi) x = hi;
    return (uint32_t)x;
}

}  // namespace ab

using namespace ab;

extern "C" ab_status ab_gen_pairs(uint64_t seed, uint64_t perm_seed, uint32_t nvalues, uint32_t nkeys,
                                  int len_dist, double p1, double p2, uint32_t lo, uint32_t hi,
                                  double zipf_s, int nthreads, ab_dataset **out)
{
    if (!out || nkeys == 0 || lo > hi || lo == 0 || (len_dist != 0 && len_dist != 1)) {
        set_error("ab_gen_pairs: invalid arguments");
        return AB_INVALID_ARG;
    }
    if (nthreads <= 0) nthreads = default_threads();
    ZipfAlias z(nkeys, zipf_s, perm_seed);
    std::vector<std::vector<uint32_t>> pk(nthreads), pv(nthreads);
    parallel_ranges(nvalues, nthreads, [&](uint64_t r0, uint64_t r1, int t) {
        std::vector<uint32_t> row;
        for (uint64_t v = r0; v < r1; ++v) {
            Rng g(stream_seed(seed, v, 1));
            uint32_t L = draw_len(g, len_dist, p1, p2, lo, hi);
            if (L > nkeys) L = nkeys;
            row.clear();
            uint64_t tries = 0;
            while (row.size() < L) {
                uint32_t k = (++tries < 1000ull * L) ? z.sample(g) : g.below(nkeys);
                if (std::find(row.begin(), row.end(), k) == row.end()) row.push_back(k);
            }
            for (uint32_t k : row) {
                pk[t].push_back(k);
                pv[t].push_back((uint32_t)v);
            }
        }
    });
    auto *d = new ab_dataset();
    uint64_t total = 0;
    for (int t = 0; t < nthreads; ++t) total += pk[t].size();
    d->a.reserve(total);
    d->b.reserve(total);
    for (int t = 0; t < nthreads; ++t) {
        d->a.insert(d->a.end(), pk[t].begin(), pk[t].end());
        d->b.insert(d->b.end(), pv[t].begin(), pv[t].end());
        std::vector<uint32_t>().swap(pk[t]);
        std::vector<uint32_t>().swap(pv[t]);
    }
    d->count = total;
    *out = d;
    return AB_OK;
}

extern "C" ab_status ab_gen_queries(uint64_t seed, uint64_t perm_seed, uint32_t nkeys, uint32_t nq,
                                    int size_dist, double p1, double p2, uint32_t lo, uint32_t hi,
                                    int key_dist, double zipf_s, ab_dataset **out)
{
    if (!out || nkeys == 0 || lo > hi || size_dist < 0 || size_dist > 2 || key_dist < 0 || key_dist > 1 ||
        (size_dist == 2 && (lo == 0 || p1 == 1.0))) {
        set_error("ab_gen_queries: invalid arguments");
        return AB_INVALID_ARG;
    }
    std::unique_ptr<ZipfAlias> z;
    if (key_dist == 1) z.reset(new ZipfAlias(nkeys, zipf_s, perm_seed));
    auto *d = new ab_dataset();
    d->off.assign((size_t)nq + 1, 0);
    std::unordered_set<uint32_t> seen;
    for (uint32_t q = 0; q < nq; ++q) {
        Rng g(stream_seed(seed, q, 2));
        uint32_t n = size_dist == 0 ? lo : draw_len(g, size_dist, p1, p2, lo, hi);
        if (n > nkeys) n = nkeys;
        seen.clear();
        uint64_t tries = 0;
        while (seen.size() < n) {
            uint32_t k = (key_dist == 1 && ++tries < 1000ull * n + 1000) ? z->sample(g) : g.below(nkeys);
            if (seen.insert(k).second) d->a.push_back(k);
        }
        d->off[q + 1] = d->a.size();
    }
    d->count = nq;
    *out = d;
    return AB_OK;
}

extern "C" uint64_t ab_dataset_len(const ab_dataset *d) { return d ? d->a.size() : 0; }
extern "C" uint64_t ab_dataset_count(const ab_dataset *d) { return d ? d->count : 0; }
extern "C" const uint32_t *ab_dataset_a(const ab_dataset *d) { return d ? d->a.data() : nullptr; }
extern "C" const uint32_t *ab_dataset_b(const ab_dataset *d) { return d && !d->b.empty() ? d->b.data() : nullptr; }
extern "C" const uint64_t *ab_dataset_offsets(const ab_dataset *d)
{
    return d && !d->off.empty() ? d->off.data() : nullptr;
}
extern "C" void ab_dataset_free(ab_dataset *d) { delete d; }

// ---- text loaders (ReadCOO / ReadQuery, test/adsfilter_q_all.cu:127-193) ----

namespace {
bool read_file(const char *path, std::string &s)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    s.resize(n > 0 ? (size_t)n : 0);
    size_t got = n > 0 ? fread(&s[0], 1, (size_t)n, f) : 0;
    fclose(f);
    s.resize(got);
    return true;
}
inline bool ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }
// One unsigned 32-bit number after blanks; a sign or overflow fails.
inline bool num(const char *&p, const char *e, uint32_t &out)
{
    while (p < e && (*p == ' ' || *p == '\t')) ++p;
    if (p >= e || *p < '0' || *p > '9') return false;
    uint64_t v = 0;
    while (p < e && *p >= '0' && *p <= '9') {
        v = v * 10 + (uint64_t)(*p - '0');
        if (v > 0xffffffffull) return false;
        ++p;
    }
    out = (uint32_t)v;
    return true;
}
// Calls f(line_begin, line_end) for every trimmed line that is not blank or a
// '#' / '%' comment.
template <class F>
void for_lines(const char *p, const char *end, F &&f)
{
    while (p < end) {
        const char *e = (const char *)memchr(p, '\n', end - p);
        if (!e) e = end;
        const char *a = p, *b = e;
        while (a < b && ws(*a)) ++a;
        while (b > a && ws(b[-1])) --b;
        if (a < b && *a != '#' && *a != '%') f(a, b);
        p = e + 1;
    }
}
template <class F>
void for_lines(const std::string &s, F &&f)
{
    for_lines(s.data(), s.data() + s.size(), f);
}

// ReadCOO over the whole text with one thread per core: the text is cut at
// line starts into equal chunks, each parsed into its own arrays, and the
// chunks are joined in file order.
void parse_coo(const std::string &s, std::vector<uint32_t> &keys, std::vector<uint32_t> &vals, uint32_t &mk,
               uint32_t &mv)
{
    int T = s.size() < (1u << 22) ? 1 : default_threads();
    std::vector<size_t> cut(T + 1, s.size());
    cut[0] = 0;
    for (int t = 1; t < T; ++t) {
        size_t pos = std::max(cut[t - 1], s.size() / T * t);
        const char *nl = pos < s.size() ? (const char *)memchr(s.data() + pos, '\n', s.size() - pos) : nullptr;
        cut[t] = nl ? (size_t)(nl - s.data()) + 1 : s.size();
    }
    std::vector<std::vector<uint32_t>> tk(T), tv(T);
    std::vector<uint32_t> tmk(T, 0), tmv(T, 0);
    parallel_ranges(T, T, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t t = b; t < e; ++t) {
            tk[t].reserve((cut[t + 1] - cut[t]) / 8 + 16);
            tv[t].reserve((cut[t + 1] - cut[t]) / 8 + 16);
            for_lines(s.data() + cut[t], s.data() + cut[t + 1], [&](const char *a, const char *z) {
                uint32_t k, v;
                if (!num(a, z, k) || !num(a, z, v)) return;
                tk[t].push_back(k);
                tv[t].push_back(v);
                tmk[t] = std::max(tmk[t], k);
                tmv[t] = std::max(tmv[t], v);
            });
        }
    });
    size_t n = 0;
    for (int t = 0; t < T; ++t) n += tk[t].size();
    keys.resize(n);
    vals.resize(n);
    size_t o = 0;
    mk = mv = 0;
    for (int t = 0; t < T; ++t) {
        if (!tk[t].empty()) {
            memcpy(keys.data() + o, tk[t].data(), tk[t].size() * 4);
            memcpy(vals.data() + o, tv[t].data(), tv[t].size() * 4);
        }
        o += tk[t].size();
        mk = std::max(mk, tmk[t]);
        mv = std::max(mv, tmv[t]);
    }
}
}  // namespace

namespace ab {
// Parses a COO text file (ReadCOO semantics) in parallel; false if unreadable.
bool read_coo_file(const char *path, std::vector<uint32_t> &keys, std::vector<uint32_t> &vals, uint32_t &mk,
                   uint32_t &mv)
{
    std::string s;
    if (!path || !read_file(path, s)) return false;
    parse_coo(s, keys, vals, mk, mv);
    return true;
}
}  // namespace ab

extern "C" int64_t ab_read_query(const char *path, uint32_t *keys, uint64_t cap)
{
    std::string s;
    if (!path || !read_file(path, s)) {
        set_error(std::string("ab_read_query: cannot open ") + (path ? path : "(null)"));
        return -1;
    }
    int64_t n = 0;
    for_lines(s, [&](const char *a, const char *b) {
        uint32_t k;
        if (!num(a, b, k)) return;
        if (keys && (uint64_t)n < cap) keys[n] = k;
        ++n;
    });
    return n;
}

// FindQueryFiles (test/adsfilter_q_all.cu:211-231): regular files (symlinks
// followed) of dir whose whole name matches query_vec_[0-9]+\.coo, as dir
// joined with the name, sorted as byte strings.
static bool is_query_name(const char *n)
{
    static const char kPre[] = "query_vec_", kSuf[] = ".coo";
    const size_t len = strlen(n), pre = sizeof(kPre) - 1, suf = sizeof(kSuf) - 1;
    if (len < pre + 1 + suf || strncmp(n, kPre, pre) != 0 || strcmp(n + len - suf, kSuf) != 0) return false;
    for (size_t i = pre; i < len - suf; ++i)
        if (n[i] < '0' || n[i] > '9') return false;
    return true;
}

extern "C" int64_t ab_find_query_files(const char *dir, char *buf, uint64_t cap, uint64_t *need)
{
    if (need) *need = 0;
    if (!dir) {
        set_error("ab_find_query_files: NULL directory");
        return -1;
    }
    DIR *d = opendir(dir);
    if (!d) {
        set_error(std::string("ab_find_query_files: cannot read directory ") + dir);
        return -1;
    }
    std::string base(dir);
    if (!base.empty() && base.back() != '/') base += '/';
    std::vector<std::string> files;
    while (struct dirent *e = readdir(d)) {
        if (!is_query_name(e->d_name)) continue;
        std::string p = base + e->d_name;
        struct stat st;
        if (stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode)) files.push_back(std::move(p));
    }
    closedir(d);
    std::sort(files.begin(), files.end());
    uint64_t total = 0;
    for (const auto &f : files) total += f.size() + 1;
    if (need) *need = total;
    if (buf && cap >= total) {
        char *o = buf;
        for (const auto &f : files) {
            memcpy(o, f.c_str(), f.size() + 1);
            o += f.size() + 1;
        }
    }
    return (int64_t)files.size();
}

extern "C" int64_t ab_read_coo(const char *path, uint32_t *keys, uint32_t *values, uint64_t cap,
                               uint32_t *max_key, uint32_t *max_value)
{
    std::vector<uint32_t> k, v;
    uint32_t mk = 0, mv = 0;
    if (!read_coo_file(path, k, v, mk, mv)) {
        set_error(std::string("ab_read_coo: cannot open ") + (path ? path : "(null)"));
        return -1;
    }
    if (keys && values) {
        const uint64_t n = std::min<uint64_t>(cap, k.size());
        if (n) {
            memcpy(keys, k.data(), n * 4);
            memcpy(values, v.data(), n * 4);
        }
    }
    if (max_key) *max_key = mk;
    if (max_value) *max_value = mv;
    return (int64_t)k.size();
}

// CalculateStats, test/adsfilter_q_all.cu:258-307.
extern "C" void ab_stats(double *d, uint64_t n, double *o)
{
    for (int i = 0; i < 9; ++i) o[i] = 0;
    if (!n) return;
    std::sort(d, d + n);
    double s = 0;
    for (uint64_t i = 0; i < n; ++i) s += d[i];
    double mean = s / n;
    double med = (n % 2 == 0) ? (d[n / 2 - 1] + d[n / 2]) / 2.0 : d[n / 2];
    double sq = 0;
    for (uint64_t i = 0; i < n; ++i) sq += (d[i] - mean) * (d[i] - mean);
    o[0] = mean;
    o[1] = med;
    o[2] = std::sqrt(sq / n);
    o[3] = d[0];
    o[4] = d[n - 1];
    o[5] = med;
    const double ps[3] = {0.90, 0.95, 0.99};
    for (int k = 0; k < 3; ++k) {
        double pos = ps[k] * (double)(n - 1);
        uint64_t lo = (uint64_t)pos, up = lo + 1;
        o[6 + k] = up >= n ? d[n - 1] : d[lo] * (1 - (pos - lo)) + d[up] * (pos - lo);
    }
}
