// capi.hip -- C ABI of the AdaBool engine: index residency, contexts, and the
// orchestration of sparse / dense / compaction launches for single queries and
// batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "engine.h"

namespace ab {
static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
}  // namespace ab

using namespace ab;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                     \
            return e_ == hipErrorOutOfMemory ? AB_OOM : AB_DEVICE_ERROR;                       \
        }                                                                                      \
    } while (0)

#define AB_TRY(expr)                     \
    do {                                 \
        ab_status s_ = (expr);           \
        if (s_ != AB_OK) return s_;      \
    } while (0)

extern "C" const char *ab_last_error(void) { return g_err.c_str(); }

extern "C" const char *ab_status_str(int s)
{
    switch (s) {
    case AB_OK: return "AB_OK";
    case AB_INVALID_ARG: return "AB_INVALID_ARG";
    case AB_OOM: return "AB_OOM";
    case AB_DEVICE_ERROR: return "AB_DEVICE_ERROR";
    case AB_UNSUPPORTED: return "AB_UNSUPPORTED";
    case AB_IO_ERROR: return "AB_IO_ERROR";
    default: return "AB_UNKNOWN";
    }
}

// ---------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------
ab_status DevBuf::ensure(size_t bytes)
{
    if (bytes <= cap) return AB_OK;
    if (p) {
        hipError_t e = hipFree(p);
        p = nullptr;
        cap = 0;
        if (e != hipSuccess) {
            set_error(std::string("hipFree: ") + hipGetErrorString(e));
            return AB_DEVICE_ERROR;
        }
    }
    size_t want = bytes + bytes / 4;
    want = (want + 255) / 256 * 256;
    if (want < 256) want = 256;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        set_error(std::string("hipMalloc(") + std::to_string(want) + "): " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? AB_OOM : AB_DEVICE_ERROR;
    }
    cap = want;
    return AB_OK;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

ab_status PinBuf::ensure(size_t bytes)
{
    if (bytes <= cap) return AB_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        set_error(std::string("hipHostMalloc: ") + hipGetErrorString(e));
        return AB_OOM;
    }
    cap = want;
    return AB_OK;
}

void PinBuf::release()
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

// ---------------------------------------------------------------------------
// index
// ---------------------------------------------------------------------------
template <class T>
static ab_status upload_vec(ab_index *ix, const std::vector<T> &v, const T **out)
{
    size_t bytes = v.size() * sizeof(T);
    size_t alloc = (bytes + 64 + 255) / 256 * 256;  // +64: vector loads never run past the end
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, alloc));
    ix->allocs.push_back(p);
    if (bytes) HIP_TRY(hipMemcpy(p, v.data(), bytes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset((uint8_t *)p + bytes, 0, alloc - bytes));
    ix->dev_bytes += alloc;
    *out = (const T *)p;
    return AB_OK;
}

static void fill_meta(ab_index *ix, const HostIndex &h)
{
    ix->nkeys = h.nkeys;
    ix->nvalues = h.nvalues;
    ix->nnz = h.nnz;
    ix->R = 0;
    for (int t = 0; t < kTypes; ++t) {
        ix->nrec[t] = h.nrec(t);
        ix->R += ix->nrec[t];
    }
    ix->list_bytes = h.list_bytes();
    ix->tile_shift = h.tile_shift;
    ix->ntiles = h.ntiles;
    ix->nblocks = h.vblk.size();
    ix->kbytes = h.vkeys.size() >= 64 ? h.vkeys.size() - 64 : 0;
    ix->stride = stride_bytes(h.nvalues);
}

// Makes a built host layout resident (or keeps it on the host for host-only
// handles).
static ab_status finish_index(std::unique_ptr<ab_index> &ix, bool host_only)
{
    fill_meta(ix.get(), ix->host);
    ix->seg8 = ix->host.seg8;
    ix->host_only = host_only;
    if (!host_only) {
        HostIndex &h = ix->host;
        HIP_TRY(hipGetDevice(&ix->device));
        DevIndex &d = ix->dev;
        AB_TRY(upload_vec(ix.get(), h.seg8, &d.seg8));
        for (int t = 0; t < kTypes; ++t) {
            AB_TRY(upload_vec(ix.get(), h.hdr[t], &d.hdr[t]));
            AB_TRY(upload_vec(ix.get(), h.pay[t], &d.pay[t]));
            AB_TRY(upload_vec(ix.get(), h.vhdr[t], &d.vhdr[t]));
            AB_TRY(upload_vec(ix.get(), h.vpay[t], &d.vpay[t]));
        }
        const uint32_t *vb = nullptr;
        std::vector<uint32_t> blk(h.vblk.size() * 4);
        if (!h.vblk.empty()) memcpy(blk.data(), h.vblk.data(), blk.size() * 4);
        AB_TRY(upload_vec(ix.get(), blk, &vb));
        d.vblk = reinterpret_cast<const uint4 *>(vb);
        AB_TRY(upload_vec(ix.get(), h.vkeys, &d.vkeys));
        AB_TRY(upload_vec(ix.get(), h.tile_blk, &d.tile_blk));
        d.nkeys = h.nkeys;
        d.nvalues = h.nvalues;
        d.tile_shift = h.tile_shift;
        d.ntiles = h.ntiles;
        // only seg8 (already copied to ix->seg8) is needed on the host once resident
        h = HostIndex();
    }
    return AB_OK;
}

static ab_status do_build(const uint32_t *keys, const uint32_t *vals, uint64_t nnz, uint32_t nkeys,
                          uint32_t nvalues, uint32_t tile_shift, bool host_only, ab_index **out)
{
    if (!out) {
        set_error("ab_index_build: NULL out");
        return AB_INVALID_ARG;
    }
    *out = nullptr;
    std::unique_ptr<ab_index> ix(new ab_index());
    ab_status st = build_host_index(keys, vals, nnz, nkeys, nvalues, tile_shift, 0, &ix->host);
    if (st != AB_OK) return st;
    AB_TRY(finish_index(ix, host_only));
    *out = ix.release();
    return AB_OK;
}

// ---------------------------------------------------------------------------
// binary index file: a header, then every array of both layouts in a fixed
// order (each as u64 byte count + bytes, padded to 8), then a 64-bit FNV-1a
// checksum of everything before it.  Little-endian, as the device.
// ---------------------------------------------------------------------------
namespace {
constexpr char kMagic[8] = {'A', 'B', 'I', 'D', 'X', '0', '0', '1'};
struct FileHeader {
    char magic[8];
    uint32_t version, nkeys, nvalues, tile_shift, ntiles, nsections;
    uint64_t nnz;
};

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void *p, size_t n)
    {
        const uint8_t *b = (const uint8_t *)p;
        for (size_t i = 0; i < n; ++i) {
            h ^= b[i];
            h *= 1099511628211ull;
        }
    }
};

// The arrays of a layout, in file order.
template <class F> void for_sections(HostIndex &h, F f)
{
    f(h.seg8);
    for (int t = 0; t < kTypes; ++t) {
        f(h.hdr[t]);
        f(h.pay[t]);
        f(h.vhdr[t]);
        f(h.vpay[t]);
    }
    f(h.vblk);
    f(h.vkeys);
    f(h.tile_blk);
}

template <class T> ab_status download(const T *d, size_t n, std::vector<T> &v)
{
    v.resize(n);
    if (n) HIP_TRY(hipMemcpy(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
    return AB_OK;
}
}  // namespace

extern "C" ab_status ab_index_save(const ab_index *ix, const char *path)
{
    if (!ix || !path) {
        set_error("ab_index_save: NULL argument");
        return AB_INVALID_ARG;
    }
    HostIndex copy;
    HostIndex *h = nullptr;
    if (ix->host_only) {
        h = const_cast<HostIndex *>(&ix->host);
    } else {
        // the resident layout is the only copy: bring it back
        const DevIndex &d = ix->dev;
        copy.nkeys = ix->nkeys;
        copy.nvalues = ix->nvalues;
        copy.nnz = ix->nnz;
        copy.tile_shift = ix->tile_shift;
        copy.ntiles = ix->ntiles;
        copy.seg8 = ix->seg8;
        for (int t = 0; t < kTypes; ++t) {
            AB_TRY(download(d.hdr[t], ix->nrec[t], copy.hdr[t]));
            AB_TRY(download(d.pay[t], ix->nrec[t] * kPayloadBytes[t], copy.pay[t]));
            AB_TRY(download(d.vhdr[t], ix->nrec[t], copy.vhdr[t]));
            AB_TRY(download(d.vpay[t], ix->nrec[t] * kPayloadBytes[t], copy.vpay[t]));
        }
        std::vector<uint4> blk;
        AB_TRY(download(d.vblk, ix->nblocks, blk));
        copy.vblk.resize(blk.size());
        if (!blk.empty()) memcpy(copy.vblk.data(), blk.data(), blk.size() * sizeof(uint4));
        AB_TRY(download(d.vkeys, ix->kbytes + 64, copy.vkeys));
        AB_TRY(download(d.tile_blk, (size_t)ix->ntiles + 1, copy.tile_blk));
        h = &copy;
    }
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error(std::string("ab_index_save: cannot create ") + path);
        return AB_IO_ERROR;
    }
    Fnv sum;
    bool ok = true;
    auto put = [&](const void *p, size_t n) {
        if (n && fwrite(p, 1, n, f) != n) ok = false;
        sum.add(p, n);
    };
    FileHeader hd{};
    memcpy(hd.magic, kMagic, 8);
    hd.version = 2;
    hd.nkeys = h->nkeys;
    hd.nvalues = h->nvalues;
    hd.tile_shift = h->tile_shift;
    hd.ntiles = h->ntiles;
    hd.nsections = 1 + 4 * kTypes + 3;
    hd.nnz = h->nnz;
    put(&hd, sizeof(hd));
    const uint8_t zeros[8] = {0};
    for_sections(*h, [&](auto &v) {
        const uint64_t bytes = v.size() * sizeof(v[0]);
        put(&bytes, 8);
        put(v.data(), bytes);
        put(zeros, (8 - bytes % 8) % 8);
    });
    const uint64_t c = sum.h;
    if (ok && fwrite(&c, 8, 1, f) != 1) ok = false;
    if (fclose(f) != 0) ok = false;
    if (!ok) {
        set_error(std::string("ab_index_save: write failed: ") + path);
        return AB_IO_ERROR;
    }
    return AB_OK;
}

extern "C" ab_status ab_index_load(const char *path, int host_only, ab_index **out)
{
    if (!path || !out) {
        set_error("ab_index_load: NULL argument");
        return AB_INVALID_ARG;
    }
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error(std::string("ab_index_load: cannot open ") + path);
        return AB_IO_ERROR;
    }
    std::unique_ptr<ab_index> ix(new ab_index());
    HostIndex &h = ix->host;
    Fnv sum;
    bool ok = true;
    auto get = [&](void *p, size_t n) {
        if (n && fread(p, 1, n, f) != n) ok = false;
        if (ok) sum.add(p, n);
    };
    FileHeader hd{};
    get(&hd, sizeof(hd));
    if (!ok || memcmp(hd.magic, kMagic, 8) != 0 || hd.version != 2 || hd.nsections != 1 + 4 * kTypes + 3) {
        fclose(f);
        set_error(std::string("ab_index_load: not an index file of this version: ") + path);
        return AB_IO_ERROR;
    }
    h.nkeys = hd.nkeys;
    h.nvalues = hd.nvalues;
    h.nnz = hd.nnz;
    h.tile_shift = hd.tile_shift;
    h.ntiles = hd.ntiles;
    // section sizes are bounded by what is left of the file
    long here = ftell(f);
    fseek(f, 0, SEEK_END);
    const uint64_t fsize = (uint64_t)ftell(f);
    fseek(f, here, SEEK_SET);
    uint8_t pad[8];
    for_sections(h, [&](auto &v) {
        if (!ok) return;
        uint64_t bytes = 0;
        get(&bytes, 8);
        if (!ok || bytes % sizeof(v[0]) != 0 || bytes > fsize) {
            ok = false;
            return;
        }
        v.resize(bytes / sizeof(v[0]));
        get(v.data(), bytes);
        get(pad, (8 - bytes % 8) % 8);
    });
    uint64_t c = 0;
    const uint64_t want = sum.h;
    if (ok && fread(&c, 8, 1, f) != 1) ok = false;
    fclose(f);
    if (!ok || c != want) {
        set_error(std::string("ab_index_load: truncated or corrupt file: ") + path);
        return AB_IO_ERROR;
    }
    // the checksum only detects accidents: every invariant the kernels rely
    // on is checked as well, so a malformed file never reaches them
    std::string why;
    if (!validate_index(h, why)) {
        set_error(std::string("ab_index_load: inconsistent index file (") + why + "): " + path);
        return AB_IO_ERROR;
    }
    AB_TRY(finish_index(ix, host_only != 0));
    *out = ix.release();
    return AB_OK;
}

extern "C" ab_status ab_index_build(const uint32_t *keys, const uint32_t *vals, uint64_t nnz, uint32_t nkeys,
                                    uint32_t nvalues, uint32_t tile_shift, ab_index **out)
{
    return do_build(keys, vals, nnz, nkeys, nvalues, tile_shift, false, out);
}

extern "C" ab_status ab_index_build_host(const uint32_t *keys, const uint32_t *vals, uint64_t nnz, uint32_t nkeys,
                                         uint32_t nvalues, uint32_t tile_shift, ab_index **out)
{
    return do_build(keys, vals, nnz, nkeys, nvalues, tile_shift, true, out);
}

extern "C" ab_status ab_index_load_coo(const char *path, uint32_t tile_shift, ab_index **out)
{
    std::vector<uint32_t> k, v;
    uint32_t mk = 0, mv = 0;
    if (!read_coo_file(path, k, v, mk, mv)) {
        set_error(std::string("ab_index_load_coo: cannot open ") + (path ? path : "(null)"));
        return AB_IO_ERROR;
    }
    return ab_index_build(k.data(), v.data(), (uint64_t)k.size(), 0, 0, tile_shift, out);
}

extern "C" ab_status ab_index_free(ab_index *ix)
{
    if (!ix) return AB_OK;
    for (void *p : ix->allocs) (void)hipFree(p);
    delete ix;
    return AB_OK;
}

extern "C" ab_status ab_index_get_info(const ab_index *ix, ab_index_info *o)
{
    if (!ix || !o) {
        set_error("ab_index_get_info: NULL argument");
        return AB_INVALID_ARG;
    }
    memset(o, 0, sizeof(*o));
    o->nkeys = ix->nkeys;
    o->nvalues = ix->nvalues;
    o->nnz = ix->nnz;
    for (int t = 0; t < kTypes; ++t) o->nrec[t] = ix->nrec[t];
    o->nrec_total = ix->R;
    o->list_bytes = ix->list_bytes;
    o->tile_shift = ix->tile_shift;
    o->ntiles = ix->ntiles;
    o->nblocks = ix->nblocks;
    o->key_stream_bytes = ix->kbytes;
    o->device_bytes = ix->dev_bytes;
    o->stride_bytes = ix->stride;
    return AB_OK;
}

extern "C" uint64_t ab_result_stride_bytes(const ab_index *ix) { return ix ? ix->stride : 0; }

extern "C" ab_status ab_index_export_type(const ab_index *ix, int t, uint32_t *h_headers, uint8_t *h_payload,
                                          uint32_t *h_segs)
{
    if (!ix || t < 0 || t >= kTypes) {
        set_error("ab_index_export_type: bad handle or type");
        return AB_INVALID_ARG;
    }
    const uint64_t n = ix->nrec[t];
    if (ix->host_only) {
        if (h_headers && n) memcpy(h_headers, ix->host.hdr[t].data(), n * 4);
        if (h_payload && n && kPayloadBytes[t]) memcpy(h_payload, ix->host.pay[t].data(), n * kPayloadBytes[t]);
    } else {
        if (h_headers && n) HIP_TRY(hipMemcpy(h_headers, ix->dev.hdr[t], n * 4, hipMemcpyDeviceToHost));
        if (h_payload && n && kPayloadBytes[t])
            HIP_TRY(hipMemcpy(h_payload, ix->dev.pay[t], n * kPayloadBytes[t], hipMemcpyDeviceToHost));
    }
    if (h_segs)
        for (uint64_t k = 0; k <= ix->nkeys; ++k) h_segs[k] = ix->seg8[k * kSegStride + t];
    return AB_OK;
}

extern "C" int64_t ab_index_decode_tiles(const ab_index *ix, uint32_t *ok, uint32_t *ov, uint64_t cap)
{
    if (!ix) {
        set_error("ab_index_decode_tiles: NULL handle");
        return -1;
    }
    HostIndex tmp;
    const HostIndex *h = &ix->host;
    if (!ix->host_only) {
        // pull the value-tiled arrays back from the device
        tmp.vblk.resize(ix->nblocks);
        tmp.vkeys.resize(ix->kbytes + 64);
        tmp.tile_blk.resize(ix->ntiles + 1);
        if (ix->nblocks && hipMemcpy(tmp.vblk.data(), ix->dev.vblk, ix->nblocks * 16, hipMemcpyDeviceToHost) != hipSuccess)
            return -1;
        if (hipMemcpy(tmp.vkeys.data(), ix->dev.vkeys, tmp.vkeys.size(), hipMemcpyDeviceToHost) != hipSuccess) return -1;
        if (hipMemcpy(tmp.tile_blk.data(), ix->dev.tile_blk, tmp.tile_blk.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return -1;
        for (int t = 0; t < kTypes; ++t) {
            tmp.vhdr[t].resize(ix->nrec[t]);
            tmp.vpay[t].resize(ix->nrec[t] * kPayloadBytes[t]);
            if (ix->nrec[t] &&
                hipMemcpy(tmp.vhdr[t].data(), ix->dev.vhdr[t], ix->nrec[t] * 4, hipMemcpyDeviceToHost) != hipSuccess)
                return -1;
            if (!tmp.vpay[t].empty() &&
                hipMemcpy(tmp.vpay[t].data(), ix->dev.vpay[t], tmp.vpay[t].size(), hipMemcpyDeviceToHost) != hipSuccess)
                return -1;
        }
        h = &tmp;
    }
    uint64_t c = 0;
    auto put = [&](uint32_t k, uint32_t v) {
        if (ok && ov && c < cap) {
            ok[c] = k;
            ov[c] = v;
        }
        ++c;
    };
    for (uint64_t b = 0; b < h->vblk.size(); ++b) {
        const BlockMeta &m = h->vblk[b];
        const uint32_t t = m.tw & 15, wl = (m.tw >> 4) & 3, nv = m.tw >> 8, w = 1u << wl;
        for (uint32_t r = 0; r < nv; ++r) {
            const uint32_t lane = r & 63, slot = r >> 6;
            const uint8_t *p = &h->vkeys[((uint64_t)m.off_kb << 10) + ((uint64_t)lane * kSlots + slot) * w];
            uint32_t d = 0;
            for (uint32_t q = 0; q < w; ++q) d |= (uint32_t)p[q] << (8 * q);
            const uint32_t key = m.start_key + d;
            const uint64_t idx = (uint64_t)m.rec_base + r;
            const uint32_t hd = h->vhdr[t][idx];
            const uint8_t *pl = h->vpay[t].data() + idx * kPayloadBytes[t];
            if (t == AB_BITSET) {
                for (int x = 0; x < 256; ++x)
                    if (pl[x >> 3] & (1u << (x & 7))) put(key, (hd & 0xffffff00u) | (uint32_t)x);
            } else {
                put(key, hd);
                for (int i = 0; i < kPayloadBytes[t]; ++i) {
                    if (i >= kMinNum[t] && pl[i] == 0) break;
                    put(key, (hd & 0xffffff00u) | pl[i]);
                }
            }
        }
    }
    return (int64_t)c;
}

// ComputeKeyBlockMeta (src/AdaBool/inverted_list.h:90-148) for the reference
// layout: blocks of `step` keys, width by span to the next block start.
extern "C" uint64_t ab_key_block_meta(uint32_t step, uint64_t len, const uint32_t *keys, uint32_t *starts,
                                      uint32_t *pos, uint8_t *buf)
{
    if (!len || !step || !keys) return 0;
    const uint64_t nb = (len + step - 1) / step;
    uint64_t bytes = 0;
    uint32_t p = 0;
    if (pos) pos[0] = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        const uint32_t beg = keys[b * step];
        const uint32_t nxt = b + 1 < nb ? keys[(b + 1) * step] : keys[len - 1];
        const uint32_t span = nxt - beg;
        const uint32_t w = span < 256u ? 1 : (span < 65536u ? 2 : 4);
        const uint64_t num = std::min<uint64_t>(step, len - b * step);
        if (starts) starts[b] = beg;
        p += (uint32_t)(num * w);
        if (pos) pos[b + 1] = p;
        if (buf) {
            for (uint64_t i = 0; i < num; ++i) {
                const uint32_t d = keys[b * step + i] - beg;
                for (uint32_t j = 0; j < w; ++j) buf[bytes + i * w + j] = (uint8_t)(d >> (8 * j));
            }
            memset(buf + bytes + num * w, 0, (step - num) * w);
        }
        bytes += (uint64_t)step * w;
    }
    if (starts) starts[nb] = keys[len - 1];
    return bytes;
}

extern "C" ab_status ab_query_features_host(const ab_index *ix, const uint32_t *keys, uint32_t n, float *f,
                                            uint64_t *m)
{
    if (!ix || !f || (n && !keys)) {
        set_error("ab_query_features_host: NULL argument");
        return AB_INVALID_ARG;
    }
    host_features(ix->seg8.data(), ix->nkeys, ix->nvalues, ix->R, keys, n, f, m);
    return AB_OK;
}

// ---------------------------------------------------------------------------
// contexts
// ---------------------------------------------------------------------------
extern "C" ab_status ab_ctx_create(ab_index *ix, void *stream, ab_ctx **out)
{
    if (!ix || !out) {
        set_error("ab_ctx_create: NULL argument");
        return AB_INVALID_ARG;
    }
    if (ix->host_only) {
        set_error("ab_ctx_create: index was built host-only");
        return AB_INVALID_ARG;
    }
    std::unique_ptr<ab_ctx> c(new ab_ctx());
    c->ix = ix;
    if (stream) {
        c->stream = (hipStream_t)stream;
        c->own_stream = false;
    } else {
        HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
    }
    HIP_TRY(hipEventCreateWithFlags(&c->stage_ev, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&c->kstage_ev, hipEventDisableTiming));
    // clean query bitmap (1 bit per key) and query masks (u64 per key)
    AB_TRY(c->qbm.ensure(((uint64_t)ix->nkeys + 31) / 32 * 4 + 64));
    HIP_TRY(hipMemsetAsync(c->qbm.p, 0, c->qbm.cap, c->stream));
    // set-up of ab_list_union_sparse_noalloc on this device; a failure is
    // kept and reported by that entry point, not by context creation
    {
        const std::string keep = g_err;
        if (noalloc_setup(ix->device) != AB_OK) g_err = keep;
    }
    *out = c.release();
    return AB_OK;
}

ab_ctx::~ab_ctx()
{
    release_graph(this);
    for (DevBuf *b : {&qoff, &sel, &kq, &cnt, &koff, &parts, &feat, &M, &ccnt, &bm, &qlist, &keys, &qbm, &mask,
                      &mlp[0].params, &mlp[0].norm, &mlp[1].params, &mlp[1].norm, &qgroup, &grp_q, &grp_n, &plan,
                      &counter, &gmask, &ganybm, &idoff, &id_vals, &id_sorted, &id_ccnt, &id_cbase, &id_start,
                      &id_tmp, &gkeys, &sp_flag, &sp_cpre, &sp_scpre, &sp_rbase, &sp_slist, &sp_tot, &sp_refs,
                      &sp_choff, &idpart})
        b->release();
    for (PinBuf *b : {&hstage, &hkstage, &hsel, &hM, &hidpart})
        b->release();
    if (stage_ev) (void)hipEventDestroy(stage_ev);
    if (kstage_ev) (void)hipEventDestroy(kstage_ev);
    for (auto &sp : spans) {
        (void)hipEventDestroy(sp.a);
        (void)hipEventDestroy(sp.b);
    }
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}

extern "C" ab_status ab_ctx_free(ab_ctx *c)
{
    if (!c) return AB_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;
    return AB_OK;
}

extern "C" ab_status ab_ctx_sync(ab_ctx *c)
{
    if (!c) {
        set_error("ab_ctx_sync: NULL context");
        return AB_INVALID_ARG;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return AB_OK;
}

extern "C" void *ab_ctx_stream(ab_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" ab_status ab_ctx_set_threshold(ab_ctx *c, double tau)
{
    if (!c || !(tau >= 0)) {
        set_error("ab_ctx_set_threshold: bad argument");
        return AB_INVALID_ARG;
    }
    c->tau = tau;
    c->cfg_gen++;
    return AB_OK;
}

extern "C" double ab_ctx_get_threshold(const ab_ctx *c) { return c ? c->tau : 0; }

extern "C" ab_status ab_ctx_set_batch_threshold(ab_ctx *c, double tau)
{
    if (!c || !(tau >= 0)) {
        set_error("ab_ctx_set_batch_threshold: bad argument");
        return AB_INVALID_ARG;
    }
    c->tau_batch = tau;
    c->cfg_gen++;
    return AB_OK;
}

extern "C" double ab_ctx_get_batch_threshold(const ab_ctx *c) { return c ? c->tau_batch : 0; }

extern "C" ab_status ab_ctx_set_dense_group(ab_ctx *c, uint32_t w)
{
    if (!c || w < 1 || w > kGroupSlots) {
        set_error("ab_ctx_set_dense_group: width must be 1..256");
        return AB_INVALID_ARG;
    }
    c->group = w;
    c->cfg_gen++;
    return AB_OK;
}

extern "C" ab_status ab_ctx_set_dense_threads(ab_ctx *c, uint32_t t)
{
    if (!c || (t != 64 && t != 128 && t != 256 && t != 512)) {
        set_error("ab_ctx_set_dense_threads: 64, 128, 256 or 512");
        return AB_INVALID_ARG;
    }
    c->dense_threads = t;
    c->cfg_gen++;
    return AB_OK;
}

extern "C" ab_status ab_ctx_set_dense_columns(ab_ctx *c, int on)
{
    if (!c) {
        set_error("ab_ctx_set_dense_columns: NULL context");
        return AB_INVALID_ARG;
    }
    c->dense_columns = on != 0;
    c->cfg_gen++;
    return AB_OK;
}

extern "C" ab_status ab_ctx_set_batch_tuning(ab_ctx *c, const ab_tuning *e, uint32_t n)
{
    if (!c || (n && !e)) {
        set_error("ab_ctx_set_batch_tuning: bad argument");
        return AB_INVALID_ARG;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const ab_tuning &t = e[i];
        const bool ok = t.min_batch >= 1 && (i == 0 || t.min_batch > e[i - 1].min_batch) && t.group >= 1 &&
                        t.group <= kGroupSlots &&
                        (t.dense_threads == 64 || t.dense_threads == 128 || t.dense_threads == 256 ||
                         t.dense_threads == 512) &&
                        t.sparse_chunk >= 256 && t.sparse_chunk <= 8192 && !(t.sparse_chunk & (t.sparse_chunk - 1));
        if (!ok) {
            set_error("ab_ctx_set_batch_tuning: entry " + std::to_string(i) +
                      " out of range or min_batch not strictly ascending");
            return AB_INVALID_ARG;
        }
    }
    c->tuning.assign(e, e + n);
    return AB_OK;
}

extern "C" ab_status ab_ctx_get_batch_settings(const ab_ctx *c, ab_tuning *out)
{
    if (!c || !out) {
        set_error("ab_ctx_get_batch_settings: NULL argument");
        return AB_INVALID_ARG;
    }
    *out = ab_tuning{0, c->group, c->dense_threads, c->sparse_chunk, c->dense_columns ? 1 : 0};
    return AB_OK;
}

// The tuning table's entry for a batch of nq queries, applied (settings that
// change bump the launch generation).
static void apply_batch_tuning(ab_ctx *c, uint32_t nq)
{
    const ab_tuning *pick = nullptr;
    for (const ab_tuning &t : c->tuning)
        if (t.min_batch <= nq) pick = &t;
    if (!pick) return;
    const bool col = pick->dense_columns != 0;
    if (c->group == pick->group && c->dense_threads == pick->dense_threads &&
        c->sparse_chunk == pick->sparse_chunk && c->dense_columns == col)
        return;
    c->group = pick->group;
    c->dense_threads = pick->dense_threads;
    c->sparse_chunk = pick->sparse_chunk;
    c->dense_columns = col;
    c->cfg_gen++;
}

extern "C" ab_status ab_ctx_set_batch_mode(ab_ctx *c, int mode)
{
    if (!c || mode < 0 || mode > 2) {
        set_error("ab_ctx_set_batch_mode: 0, 1 or 2");
        return AB_INVALID_ARG;
    }
    c->batch_mode = mode;
    c->cfg_gen++;
    return AB_OK;
}

extern "C" ab_status ab_ctx_set_sparse_chunk(ab_ctx *c, uint32_t items)
{
    if (!c || items < 256 || items > 8192 || (items & (items - 1))) {
        set_error("ab_ctx_set_sparse_chunk: power of two in 256..8192");
        return AB_INVALID_ARG;
    }
    c->sparse_chunk = items;
    c->cfg_gen++;
    return AB_OK;
}

// ---------------------------------------------------------------------------
// per-kernel-class timing
// ---------------------------------------------------------------------------
static hipEvent_t take_event(ab_ctx *c)
{
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimedSpan::TimedSpan(ab_ctx *ctx, int k) : c(ctx), kind(k)
{
    if (!c->timing) return;
    a = take_event(c);
    if (a) (void)hipEventRecord(a, c->stream);
}

TimedSpan::~TimedSpan()
{
    if (!c->timing || !a) return;
    hipEvent_t b = take_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->spans.push_back({kind, a, b});
}

extern "C" ab_status ab_ctx_set_timing(ab_ctx *c, int on)
{
    if (!c) {
        set_error("ab_ctx_set_timing: NULL context");
        return AB_INVALID_ARG;
    }
    c->timing = on != 0;
    return AB_OK;
}

extern "C" ab_status ab_ctx_timing(ab_ctx *c, double *ms, uint64_t *n, int reset)
{
    if (!c) {
        set_error("ab_ctx_timing: NULL context");
        return AB_INVALID_ARG;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (auto &s : c->spans) {
        float e = 0;
        HIP_TRY(hipEventElapsedTime(&e, s.a, s.b));
        c->t_ms[s.kind] += e;
        c->t_n[s.kind] += 1;
        c->ev_pool.push_back(s.a);
        c->ev_pool.push_back(s.b);
    }
    c->spans.clear();
    for (int i = 0; i < AB_K_NCLASSES; ++i) {
        if (ms) ms[i] = c->t_ms[i];
        if (n) n[i] = c->t_n[i];
        if (reset) {
            c->t_ms[i] = 0;
            c->t_n[i] = 0;
        }
    }
    return AB_OK;
}

// ---------------------------------------------------------------------------
// orchestration
// ---------------------------------------------------------------------------
int ab::group_mode(uint32_t B)
{
    return B == 1 ? kModeBitmap : (B <= 32 ? kModeMask32 : (B <= 64 ? kModeMask64 : (B <= 128 ? kModeMask128 : kModeMask256)));
}

constexpr uint32_t kLdsBudget = 65536;

// Largest dense group whose LDS (key staging and row tiles) fits the dynamic
// LDS budget; the column masks are added when they fit too.
uint32_t ab::max_group(const ab_ctx *c)
{
    uint32_t g = c->group;
    while (g > 1 && dense_lds_bytes(group_mode(g), g, c->ix->tile_shift, c->dense_threads, false) > kLdsBudget) --g;
    return g;
}

bool ab::col_fits(const ab_ctx *c, int mode, uint32_t B)
{
    return c->dense_columns && dense_lds_bytes(mode, B, c->ix->tile_shift, c->dense_threads, true) <= kLdsBudget;
}

// Waits until the pinned staging area is no longer read by an earlier copy.
ab_status ab::stage_wait(ab_ctx *c)
{
    if (c->stage_pending) {
        HIP_TRY(hipEventSynchronize(c->stage_ev));
        c->stage_pending = false;
    }
    return AB_OK;
}

ab_status ab::run_batch(ab_ctx *c, const BatchArgs &a)
{
    ab_index *ix = c->ix;
    const DevIndex &d = ix->dev;
    hipStream_t s = c->stream;
    const uint32_t nq = a.nq;
    if (nq == 0) return AB_OK;
    const uint64_t K = a.h_qoff[nq];
    const uint64_t out_words = a.out_words, out_stride = a.out_stride;

    // ---- stage qoff (+ host decisions, query lists) through pinned memory -----
    // layout of hstage: qoff [nq+1 u64] | sel [nq u8, padded to 8] | sparse list [nq u32, padded to 8]
    //                   | id offsets [nq u64]
    AB_TRY(stage_wait(c));
    const size_t qoff_bytes = ((size_t)nq + 1) * 8;
    const size_t sel_off = qoff_bytes, list_off = qoff_bytes + ((size_t)nq + 7) / 8 * 8;
    const size_t idoff_off = list_off + ((size_t)nq * 4 + 7) / 8 * 8;
    AB_TRY(c->hstage.ensure(idoff_off + (size_t)nq * 8 + 64));
    memcpy(c->hstage.p, a.h_qoff, qoff_bytes);
    AB_TRY(c->qoff.ensure(qoff_bytes));
    HIP_TRY(hipMemcpyAsync(c->qoff.p, c->hstage.p, qoff_bytes, hipMemcpyHostToDevice, s));
    const uint64_t *d_idoff = nullptr;
    if (a.h_id_off && a.d_ids) {
        uint8_t *st = (uint8_t *)c->hstage.p + idoff_off;
        memcpy(st, a.h_id_off, (size_t)nq * 8);
        AB_TRY(c->idoff.ensure((size_t)nq * 8));
        HIP_TRY(hipMemcpyAsync(c->idoff.p, st, (size_t)nq * 8, hipMemcpyHostToDevice, s));
        d_idoff = (const uint64_t *)c->idoff.p;
    }
    AB_TRY(c->sel.ensure(nq));
    AB_TRY(c->hsel.ensure(nq));
    AB_TRY(c->hM.ensure((size_t)nq * 8));
    uint8_t *hsel = (uint8_t *)c->hsel.p;
    uint64_t *hM = (uint64_t *)c->hM.p;
    bool have_M = false;
    const uint64_t *qoff_d = (const uint64_t *)c->qoff.p;

    if (a.h_sel) {
        memcpy(hsel, a.h_sel, nq);
        uint8_t *st = (uint8_t *)c->hstage.p + sel_off;
        memcpy(st, a.h_sel, nq);
        HIP_TRY(hipMemcpyAsync(c->sel.p, st, nq, hipMemcpyHostToDevice, s));
        if (a.h_M) {
            memcpy(hM, a.h_M, (size_t)nq * 8);
            have_M = true;
        }
    } else if (a.path == AB_PATH_DENSE || a.path == AB_PATH_SPARSE) {
        memset(hsel, a.path, nq);
        uint8_t *st = (uint8_t *)c->hstage.p + sel_off;
        memset(st, a.path, nq);
        HIP_TRY(hipMemcpyAsync(c->sel.p, st, nq, hipMemcpyHostToDevice, s));
    } else {
        // device-side features and decision, one host round trip
        AB_TRY(c->M.ensure((size_t)nq * 8));
        AB_TRY(c->feat.ensure((size_t)nq * AB_NFEATURES * 4));
        {
            TimedSpan ts(c, AB_K_DECIDE);
            launch_features(d, a.d_keys, qoff_d, nq, ix->R, (float *)c->feat.p, (uint64_t *)c->M.p, s);
            // one dense pass serves a whole group inside a batch: batches are
            // priced by the batch model, single queries by the single model
            const int slot = nq > 1 ? 1 : 0;
            const bool learned = (a.path == AB_PATH_LEARNED || a.path == AB_PATH_AUTO) && c->mlp[slot].loaded;
            if (learned) {
                AB_TRY(mlp_decide(c, slot, (const float *)c->feat.p, nq, (uint8_t *)c->sel.p));
            } else {
                const double tau = nq > 1 ? c->tau_batch : c->tau;
                launch_decide_threshold((const uint64_t *)c->M.p, nq, tau * (double)ix->R, (uint8_t *)c->sel.p, s);
            }
        }
        HIP_TRY(hipMemcpyAsync(hsel, c->sel.p, nq, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(hM, c->M.p, (size_t)nq * 8, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        have_M = true;
    }

    std::vector<uint32_t> dq, sq;
    uint64_t Wsparse = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        if (hsel[q] == AB_PATH_DENSE) dq.push_back(q);
        else {
            sq.push_back(q);
            if (have_M) Wsparse += hM[q];
        }
    }
    if (a.h_paths) memcpy(a.h_paths, hsel, nq);

    // ---- output bitmaps --------------------------------------------------------
    uint32_t *bm = a.d_bitmaps;
    if (!bm) {
        AB_TRY(c->bm.ensure((size_t)nq * out_stride * 4));
        bm = (uint32_t *)c->bm.p;
    }

    // ---- sparse queries: zero their bitmaps, plan, scan, union ---------------
    if (!sq.empty()) {
        {
            TimedSpan ts(c, AB_K_SPARSE_ZERO);
            if (sq.size() == nq && out_stride == out_words) {
                launch_fill_zero(bm, (uint64_t)nq * out_stride * 4, s);
            } else {
                uint32_t *st = (uint32_t *)((uint8_t *)c->hstage.p + list_off);
                memcpy(st, sq.data(), sq.size() * 4);
                AB_TRY(c->qlist.ensure(sq.size() * 4));
                HIP_TRY(hipMemcpyAsync(c->qlist.p, st, sq.size() * 4, hipMemcpyHostToDevice, s));
                launch_zero_qlist(bm, out_stride, out_words, (const uint32_t *)c->qlist.p, (uint32_t)sq.size(), s);
            }
        }
        if (K > 0 && (!have_M || Wsparse > 0)) {
            AB_TRY(c->kq.ensure(K * 4));
            AB_TRY(c->cnt.ensure(K * 8));
            AB_TRY(c->koff.ensure((K + 1) * 8));
            AB_TRY(c->parts.ensure(scan_parts_needed(K) * 8 + 16));
            {
                TimedSpan ts(c, AB_K_SPARSE_PLAN);
                launch_sparse_plan(d, a.d_keys, qoff_d, (const uint8_t *)c->sel.p, nq, (uint32_t *)c->kq.p,
                                   (uint64_t *)c->cnt.p, s);
                launch_scan_u64((const uint64_t *)c->cnt.p, (uint64_t *)c->koff.p, K, (uint64_t *)c->parts.p, s);
            }
            uint32_t grid = 2048;
            if (have_M) {
                uint64_t nch = (Wsparse + c->sparse_chunk - 1) / c->sparse_chunk;
                grid = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(nch, 1), 8192);
            }
            TimedSpan ts(c, AB_K_SPARSE_UNION);
            launch_sparse_union(d, a.d_keys, (const uint32_t *)c->kq.p, (const uint64_t *)c->koff.p, K, bm,
                                out_stride, c->sparse_chunk, grid, s);
        }
    }

    // ---- dense queries: groups of up to 64 per key-stream pass ----------------
    if (!dq.empty()) {
        // host-dispatched groups carry at most 64 queries (one u64 mask word)
        const uint32_t G = std::min<uint32_t>(max_group(c), kMaxGroup);
        for (size_t g0 = 0; g0 < dq.size(); g0 += G) {
            DenseGroup g;
            g.B = (uint32_t)std::min<size_t>(G, dq.size() - g0);
            for (uint32_t b = 0; b < g.B; ++b) g.q[b] = dq[g0 + b];
            const int mode = group_mode(g.B);
            void *mask = c->qbm.p;
            if (mode != kModeBitmap) {
                if (!c->mask.p) {
                    AB_TRY(c->mask.ensure((size_t)ix->nkeys * 8 + 64));
                    HIP_TRY(hipMemsetAsync(c->mask.p, 0, c->mask.cap, s));
                }
                mask = c->mask.p;
            }
            {
                TimedSpan ts(c, AB_K_DENSE_MASK);
                launch_mask_keys(mode, a.d_keys, qoff_d, g, mask, (uint32_t *)c->qbm.p, ix->nkeys, 0, s);
            }
            {
                TimedSpan ts(c, AB_K_DENSE_TILES);
                HIP_TRY(launch_dense_tiles(mode, c->dense_threads, d, mask, (const uint32_t *)c->qbm.p, g, bm, out_stride,
                                           out_words, col_fits(c, mode, g.B), s));
            }
            TimedSpan ts(c, AB_K_DENSE_MASK);
            launch_mask_keys(mode, a.d_keys, qoff_d, g, mask, (uint32_t *)c->qbm.p, ix->nkeys, 1, s);
        }
    }

    // ---- ids / counts ------------------------------------------------------------
    if (a.d_ids || a.d_counts) {
        const uint64_t nchunk = (out_words + kCompactWords - 1) / kCompactWords;
        AB_TRY(c->ccnt.ensure((size_t)nq * nchunk * 4 + 4));
        {
            TimedSpan ts(c, AB_K_POPCOUNT);
            launch_popcount(bm, out_stride, out_words, nq, (uint32_t *)c->ccnt.p, s);
        }
        TimedSpan ts(c, AB_K_WRITE_IDS);
        launch_write_ids(bm, out_stride, out_words, nq, (uint32_t *)c->ccnt.p, a.d_ids, a.ids_stride, d_idoff,
                         a.id_off_aligned, a.d_counts, s);
    }
    HIP_TRY(hipGetLastError());
    // the staged control data stays in use until the copies above have run
    HIP_TRY(hipEventRecord(c->stage_ev, s));
    c->stage_pending = true;
    return AB_OK;
}

// Copies host keys into the context's device key buffer through pinned memory.
static ab_status upload_keys(ab_ctx *c, const uint32_t *h_keys, uint64_t n)
{
    if (c->kstage_pending) {
        HIP_TRY(hipEventSynchronize(c->kstage_ev));
        c->kstage_pending = false;
    }
    AB_TRY(c->keys.ensure(n * 4 + 16));
    if (n) {
        AB_TRY(c->hkstage.ensure(n * 4));
        memcpy(c->hkstage.p, h_keys, n * 4);
        HIP_TRY(hipMemcpyAsync(c->keys.p, c->hkstage.p, n * 4, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipEventRecord(c->kstage_ev, c->stream));
        c->kstage_pending = true;
    }
    return AB_OK;
}

// Host-side choice for one query of host keys, priced as a single query
// (batch = false) or as a share of a batch (batch = true: the batch model and
// tau_batch, as on the device).
static int choose_host(ab_ctx *c, const uint32_t *keys, uint32_t n, int path, uint64_t *M, bool batch = false)
{
    float f[AB_NFEATURES];
    host_features(c->ix->seg8.data(), c->ix->nkeys, c->ix->nvalues, c->ix->R, keys, n, f, M);
    if (path == AB_PATH_DENSE || path == AB_PATH_SPARSE) return path;
    const int slot = batch ? 1 : 0;
    if ((path == AB_PATH_LEARNED || path == AB_PATH_AUTO) && c->mlp[slot].loaded) return mlp_decide_host(c, slot, f);
    const double tau = batch ? c->tau_batch : c->tau;
    return (double)*M >= tau * (double)c->ix->R ? AB_PATH_DENSE : AB_PATH_SPARSE;
}

extern "C" ab_status ab_query(ab_ctx *c, const uint32_t *h_keys, uint32_t n, int path, uint32_t *d_bitmap,
                              uint32_t *d_ids, uint32_t *d_count, int *h_chosen)
{
    if (!c || (n && !h_keys) || path < 0 || path > AB_PATH_LEARNED) {
        set_error("ab_query: bad argument");
        return AB_INVALID_ARG;
    }
    uint64_t M = 0;
    const uint8_t sel = (uint8_t)choose_host(c, h_keys, n, path, &M);
    if (h_chosen) *h_chosen = sel;
    AB_TRY(upload_keys(c, h_keys, n));
    BatchArgs a;
    const uint64_t qoff[2] = {0, n};
    a.d_keys = (const uint32_t *)c->keys.p;
    a.h_qoff = qoff;
    a.nq = 1;
    a.h_sel = &sel;
    a.h_M = &M;
    a.d_bitmaps = d_bitmap;
    a.out_stride = a.out_words = c->ix->stride / 4;
    a.d_ids = d_ids;
    a.ids_stride = c->ix->nvalues;
    a.d_counts = d_count;
    return run_batch(c, a);
}

// A caller's result buffer of res_len bytes (at least ceil(nvalues / 8), as the
// reference's callers size it, test/adsfilter_q_all.cu:361-368): the full
// stride goes straight into it when it has room, else into the context's
// bitmap scratch, from which exactly res_len bytes are copied.
static ab_status result_target(ab_ctx *c, const char *fn, uint32_t *d_res, uint64_t res_len, uint32_t **target)
{
    const uint64_t need = ((uint64_t)c->ix->nvalues + 7) / 8;
    if (res_len < need) {
        set_error(std::string(fn) + ": res_len_bytes smaller than ceil(nvalues / 8)");
        return AB_INVALID_ARG;
    }
    *target = d_res;
    if (res_len < c->ix->stride) {
        AB_TRY(c->bm.ensure(c->ix->stride));
        *target = (uint32_t *)c->bm.p;
    }
    return AB_OK;
}

static ab_status result_finish(ab_ctx *c, uint32_t *d_res, uint64_t res_len, const uint32_t *target)
{
    if (target != d_res)
        HIP_TRY(hipMemcpyAsync(d_res, target, res_len, hipMemcpyDeviceToDevice, c->stream));
    return AB_OK;
}

extern "C" ab_status ab_list_union_sparse(ab_ctx *c, const uint32_t *h_keys, uint32_t n, uint32_t *d_res,
                                          uint64_t res_len_bytes)
{
    if (!c || (n && !h_keys) || !d_res) {
        set_error("ab_list_union_sparse: bad argument");
        return AB_INVALID_ARG;
    }
    uint32_t *t = nullptr;
    AB_TRY(result_target(c, "ab_list_union_sparse", d_res, res_len_bytes, &t));
    AB_TRY(ab_query(c, h_keys, n, AB_PATH_SPARSE, t, nullptr, nullptr, nullptr));
    return result_finish(c, d_res, res_len_bytes, t);
}

extern "C" ab_status ab_list_union_dense(ab_ctx *c, uint64_t len_bytes, const uint8_t *h_bitmap, uint32_t *d_res,
                                         uint64_t res_len_bytes)
{
    if (!c || (len_bytes && !h_bitmap) || !d_res) {
        set_error("ab_list_union_dense: bad argument");
        return AB_INVALID_ARG;
    }
    uint32_t *t = nullptr;
    AB_TRY(result_target(c, "ab_list_union_dense", d_res, res_len_bytes, &t));
    ab_index *ix = c->ix;
    const uint64_t qbytes = ((uint64_t)ix->nkeys + 31) / 32 * 4;
    const uint64_t nb = std::min<uint64_t>(len_bytes, qbytes);
    AB_TRY(stage_wait(c));
    AB_TRY(c->hstage.ensure(qbytes + 16));
    memset(c->hstage.p, 0, qbytes);
    if (nb) memcpy(c->hstage.p, h_bitmap, nb);
    // bits beyond nkeys can match no record; keep them out of the clean buffer
    HIP_TRY(hipMemcpyAsync(c->qbm.p, c->hstage.p, qbytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipEventRecord(c->stage_ev, c->stream));
    c->stage_pending = true;
    DenseGroup g;
    g.B = 1;
    g.q[0] = 0;
    const uint64_t w = ix->stride / 4;
    HIP_TRY(launch_dense_tiles(kModeBitmap, c->dense_threads, ix->dev, c->qbm.p, (const uint32_t *)c->qbm.p, g, t, w,
                               w, false, c->stream));
    HIP_TRY(hipMemsetAsync(c->qbm.p, 0, qbytes, c->stream));
    return result_finish(c, d_res, res_len_bytes, t);
}

extern "C" ab_status ab_list_union_dense_batch(ab_ctx *c, uint32_t B, uint64_t len_bytes, const uint8_t *d_bitmaps,
                                               uint32_t *d_res, uint32_t res_len_u32)
{
    if (!c || !B || (len_bytes % 4) || !d_bitmaps || !d_res) {
        set_error("ab_list_union_dense_batch: bad argument (len_bytes must be a multiple of 4)");
        return AB_INVALID_ARG;
    }
    ab_index *ix = c->ix;
    if ((uint64_t)res_len_u32 * 32 < ix->nvalues) {
        set_error("ab_list_union_dense_batch: res_len_u32 smaller than the value space");
        return AB_INVALID_ARG;
    }
    if (!c->mask.p) {
        AB_TRY(c->mask.ensure((size_t)ix->nkeys * 8 + 64));
        HIP_TRY(hipMemsetAsync(c->mask.p, 0, c->mask.cap, c->stream));
    }
    uint32_t G = 64;
    while (G > 1 && dense_lds_bytes(kModeMask64, G, ix->tile_shift, c->dense_threads, false) > kLdsBudget) --G;
    const uint64_t out_words = std::min<uint64_t>(res_len_u32, ix->stride / 4);
    for (uint32_t b0 = 0; b0 < B; b0 += G) {
        DenseGroup g;
        g.B = std::min(G, B - b0);
        for (uint32_t b = 0; b < g.B; ++b) g.q[b] = b0 + b;
        launch_mask_from_bitmaps((const uint32_t *)d_bitmaps, len_bytes / 4, b0, g.B, c->mask.p,
                                 (uint32_t *)c->qbm.p, ix->nkeys, c->stream);
        HIP_TRY(launch_dense_tiles(kModeMask64, c->dense_threads, ix->dev, c->mask.p, (const uint32_t *)c->qbm.p, g,
                                   d_res, res_len_u32, out_words, col_fits(c, kModeMask64, g.B), c->stream));
    }
    // both filters back to all-zero for the key-list paths
    HIP_TRY(hipMemsetAsync(c->mask.p, 0, (size_t)ix->nkeys * 8, c->stream));
    HIP_TRY(hipMemsetAsync(c->qbm.p, 0, c->qbm.cap, c->stream));
    return AB_OK;
}

extern "C" ab_status ab_list_union_sparse_batch(ab_ctx *c, const uint32_t *h_keys, const uint32_t *h_qidx,
                                                uint32_t n, uint32_t nq, uint32_t *d_res, uint32_t res_len_bytes)
{
    if (!c || (n && (!h_keys || !h_qidx)) || !d_res || (res_len_bytes % 4) || !nq) {
        set_error("ab_list_union_sparse_batch: bad argument");
        return AB_INVALID_ARG;
    }
    if ((uint64_t)res_len_bytes * 8 < c->ix->nvalues) {
        set_error("ab_list_union_sparse_batch: res_len_bytes smaller than the value space");
        return AB_INVALID_ARG;
    }
    // group the flattened keys by query (stable counting sort)
    std::vector<uint64_t> qoff((size_t)nq + 1, 0);
    for (uint32_t i = 0; i < n; ++i) {
        if (h_qidx[i] >= nq) {
            set_error("ab_list_union_sparse_batch: query index out of range");
            return AB_INVALID_ARG;
        }
        qoff[h_qidx[i] + 1]++;
    }
    for (uint32_t q = 0; q < nq; ++q) qoff[q + 1] += qoff[q];
    std::vector<uint32_t> keys(n);
    std::vector<uint64_t> cur(qoff.begin(), qoff.end() - 1);
    for (uint32_t i = 0; i < n; ++i) keys[cur[h_qidx[i]]++] = h_keys[i];
    AB_TRY(upload_keys(c, keys.data(), n));
    BatchArgs a;
    a.d_keys = (const uint32_t *)c->keys.p;
    a.h_qoff = qoff.data();
    a.nq = nq;
    a.path = AB_PATH_SPARSE;
    a.d_bitmaps = d_res;
    a.out_stride = res_len_bytes / 4;
    a.out_words = std::min<uint64_t>(res_len_bytes / 4, c->ix->stride / 4);
    return run_batch(c, a);
}

static ab_status query_batch(const char *fn, ab_ctx *c, const uint32_t *d_keys, const uint64_t *h_qoff, uint32_t nq,
                             int path, uint32_t *d_bitmaps, uint32_t *d_ids, uint64_t ids_stride,
                             const uint64_t *h_id_off, uint32_t *d_counts, uint8_t *h_paths)
{
    if (!c || !h_qoff || (h_qoff[nq] && !d_keys) || path < 0 || path > AB_PATH_LEARNED) {
        set_error(std::string(fn) + ": bad argument");
        return AB_INVALID_ARG;
    }
    if (d_ids && !h_id_off && ids_stride == 0) {
        set_error(std::string(fn) + ": ids_stride must be positive with d_ids");
        return AB_INVALID_ARG;
    }
    for (uint32_t q = 0; q < nq; ++q)
        if (h_qoff[q + 1] < h_qoff[q]) {
            set_error(std::string(fn) + ": offsets must be non-decreasing");
            return AB_INVALID_ARG;
        }
    apply_batch_tuning(c, nq);
    BatchArgs a;
    a.d_keys = d_keys;
    a.h_qoff = h_qoff;
    a.nq = nq;
    a.path = path;
    a.d_bitmaps = d_bitmaps;
    a.out_stride = a.out_words = c->ix->stride / 4;
    a.d_ids = d_ids;
    a.ids_stride = ids_stride;
    a.d_counts = d_counts;
    a.h_paths = h_paths;
    if (h_id_off) {
        a.h_id_off = h_id_off;
        a.id_off_aligned = true;
        for (uint32_t q = 0; q < nq; ++q) a.id_off_aligned = a.id_off_aligned && (h_id_off[q] % 4 == 0);
    }
    // batches of a few queries with few keys: the keys come back to the host
    // in one small copy and the choices are made there, as for single
    // queries; the host-dispatched launches then need no device features and
    // no second round trip (faster than a dozen device-dispatch launches)
    constexpr uint32_t kHostChoiceQueries = 4;
    constexpr uint64_t kHostChoiceKeys = 1u << 16;
    const uint64_t K = h_qoff[nq];
    if (nq >= 1 && nq <= kHostChoiceQueries && K <= kHostChoiceKeys) {
        if (c->kstage_pending) {
            HIP_TRY(hipEventSynchronize(c->kstage_ev));
            c->kstage_pending = false;
        }
        AB_TRY(c->hkstage.ensure(K * 4 + 16));
        if (K) HIP_TRY(hipMemcpyAsync(c->hkstage.p, d_keys, K * 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        const uint32_t *hk = (const uint32_t *)c->hkstage.p;
        uint8_t sel[kHostChoiceQueries];
        uint64_t M[kHostChoiceQueries];
        for (uint32_t q = 0; q < nq; ++q)
            sel[q] = (uint8_t)choose_host(c, hk + h_qoff[q], (uint32_t)(h_qoff[q + 1] - h_qoff[q]), path, &M[q],
                                          nq > 1);
        a.h_sel = sel;
        a.h_M = M;
        return run_batch(c, a);
    }
    // the persistent kernel is built for 256 and 512 threads
    if (c->batch_mode == 0 || nq == 0 || c->dense_threads < 256) return run_batch(c, a);
    return run_batch_fused(c, a);
}

extern "C" ab_status ab_query_batch(ab_ctx *c, const uint32_t *d_keys, const uint64_t *h_qoff, uint32_t nq, int path,
                                    uint32_t *d_bitmaps, uint32_t *d_ids, uint64_t ids_stride, uint32_t *d_counts,
                                    uint8_t *h_paths)
{
    return query_batch("ab_query_batch", c, d_keys, h_qoff, nq, path, d_bitmaps, d_ids, ids_stride, nullptr,
                       d_counts, h_paths);
}

extern "C" ab_status ab_query_batch_packed(ab_ctx *c, const uint32_t *d_keys, const uint64_t *h_qoff, uint32_t nq,
                                           int path, uint32_t *d_bitmaps, uint32_t *d_ids,
                                           const uint64_t *h_id_off, uint32_t *d_counts, uint8_t *h_paths)
{
    if (d_ids && !h_id_off) {
        set_error("ab_query_batch_packed: h_id_off is required with d_ids");
        return AB_INVALID_ARG;
    }
    return query_batch("ab_query_batch_packed", c, d_keys, h_qoff, nq, path, d_bitmaps, d_ids, 0,
                       d_ids ? h_id_off : nullptr, d_counts, h_paths);
}

extern "C" ab_status ab_compact_ids(ab_ctx *c, const uint32_t *d_bitmaps, uint32_t nq, uint32_t *d_ids,
                                    uint64_t ids_stride, uint32_t *d_counts)
{
    if (!c || !d_bitmaps || (!d_ids && !d_counts)) {
        set_error("ab_compact_ids: bad argument");
        return AB_INVALID_ARG;
    }
    const uint64_t w = c->ix->stride / 4;
    const uint64_t nchunk = (w + kCompactWords - 1) / kCompactWords;
    AB_TRY(c->ccnt.ensure((size_t)nq * nchunk * 4 + 4));
    launch_popcount(d_bitmaps, w, w, nq, (uint32_t *)c->ccnt.p, c->stream);
    launch_write_ids(d_bitmaps, w, w, nq, (uint32_t *)c->ccnt.p, d_ids, ids_stride, nullptr, false, d_counts,
                     c->stream);
    HIP_TRY(hipGetLastError());
    return AB_OK;
}

extern "C" ab_status ab_query_features(ab_ctx *c, const uint32_t *d_keys, const uint64_t *h_qoff, uint32_t nq,
                                       float *h_features, uint64_t *h_m)
{
    if (!c || !h_qoff || !h_features) {
        set_error("ab_query_features: bad argument");
        return AB_INVALID_ARG;
    }
    if (!nq) return AB_OK;
    AB_TRY(stage_wait(c));
    const size_t qb = ((size_t)nq + 1) * 8;
    AB_TRY(c->hstage.ensure(qb));
    memcpy(c->hstage.p, h_qoff, qb);
    AB_TRY(c->qoff.ensure(qb));
    AB_TRY(c->M.ensure((size_t)nq * 8));
    AB_TRY(c->feat.ensure((size_t)nq * AB_NFEATURES * 4));
    HIP_TRY(hipMemcpyAsync(c->qoff.p, c->hstage.p, qb, hipMemcpyHostToDevice, c->stream));
    launch_features(c->ix->dev, d_keys, (const uint64_t *)c->qoff.p, nq, c->ix->R, (float *)c->feat.p,
                    (uint64_t *)c->M.p, c->stream);
    HIP_TRY(hipMemcpyAsync(h_features, c->feat.p, (size_t)nq * AB_NFEATURES * 4, hipMemcpyDeviceToHost, c->stream));
    if (h_m) HIP_TRY(hipMemcpyAsync(h_m, c->M.p, (size_t)nq * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return AB_OK;
}
