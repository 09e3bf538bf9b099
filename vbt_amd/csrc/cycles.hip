// cycles.hip -- the reference's Cycles consumer of VBTPACK4 on the device:
// the resident blob (cycles_native_vbt/vbt_cycles_blob.{h,cpp}: a 76-byte
// header, the leaf offset table, the payload pool) built from an open asset
// without touching the host, and the two density samplers that read that
// blob straight from the packed payload:
//   standalone  cycles_native_vbt/vbt_cycles_sampler.h:165-376, the form the
//               reference's CPU probe runs: trilinear over each stream's
//               control lattice at the point's position inside its leaf;
//   kernel      blender_patch/native_cycles_vbt.patch:688-929, the form the
//               patched Cycles kernel runs: the same leaf sample at the eight
//               voxel corners around the point, mixed trilinearly; Mode2
//               leaves sample to 0.
//
// Launch shapes.  The reference gives each sample one thread.  Here the eight
// lattice corners of a stream are decoded by eight lanes of a wavefront at
// once (each runs one bracket search) and combined by lane shuffles in the
// reference's mixing order; the kernel form gives each of its eight voxel
// corners such a group, so one wavefront serves one sample.  The one-lane
// form (VBT_CYCLES_PLAIN) is the reference's structure, kept for comparison.
//
// Every float operation is IEEE single precision without contraction,
// division correctly rounded; mix(a, b, t) = a * (1 - t) + b * t with each
// operation rounded on its own.  Two cases without a defined reference result
// are fixed: reads past the blob's offset or payload sections read zero bytes,
// and a stream of resolution 0 contributes 0 (the standalone code would
// search 2^32 lattice positions and then read far past the leaf).
#include "vbt_format.h"
#include "vbt_internal.h"
#include "vbt_kernels.h"

#include <hip/hip_runtime.h>

using namespace vbtk;

namespace {

constexpr uint32_t BLOB_MAGIC = 0x31544256u;  // "VBT1" read as a little-endian word
constexpr uint32_t BLOB_HEADER_BYTES = 76u;   // 19 four-byte fields (vbt_cycles_blob.h:14-34)

struct BlobHeaderWords {
    uint32_t w[16];  // magic .. payload_byte_offset; the bbox floats stay 0
};

__global__ void k_blob_header(uint32_t* dst, BlobHeaderWords h)
{
    const uint32_t i = threadIdx.x;
    if (i < 16u) dst[i] = h.w[i];
    else if (i < 19u) dst[i] = 0u;
}

struct BlobView {
    const uint32_t* off;
    const uint32_t* pay;
    uint32_t offWords, payWords;
    uint32_t dimX, dimY, dimZ, frames, leafSize, lcx, lcy, lcz;
    int frameIndex;
    float bx, by, bz;
};

// blob_payload_view's checks (vbt_cycles_blob.cpp:89-124) and the kernel's
// magic test; sections must also sit on 4-byte boundaries.
__device__ __forceinline__ bool open_view(const uint32_t* blob, uint64_t bytes, BlobView& v)
{
    if (bytes < BLOB_HEADER_BYTES) return false;
    const uint32_t magic = blob[0], version = blob[1];
    const uint32_t offWords = blob[12], payWords = blob[13], offAt = blob[14], payAt = blob[15];
    if (magic != BLOB_MAGIC || version != 1u) return false;
    if ((uint64_t)offAt + 4ull * offWords > bytes || (uint64_t)payAt + 4ull * payWords > bytes) return false;
    if ((offAt | payAt) & 3u) return false;
    v.dimX = blob[2];
    v.dimY = blob[3];
    v.dimZ = blob[4];
    v.frames = blob[5];
    v.leafSize = blob[6];
    v.lcx = blob[7];
    v.lcy = blob[8];
    v.lcz = blob[9];
    v.frameIndex = (int)blob[11];
    v.offWords = offWords;
    v.payWords = payWords;
    v.off = blob + offAt / 4u;
    v.pay = blob + payAt / 4u;
    v.bx = __uint_as_float(blob[16]);
    v.by = __uint_as_float(blob[17]);
    v.bz = __uint_as_float(blob[18]);
    return v.leafSize != 0u;
}

__device__ __forceinline__ uint32_t pay_word(const BlobView& v, uint32_t w) { return w < v.payWords ? v.pay[w] : 0u; }

// the four payload bytes starting at byte `at` (zero past the pool)
__device__ __forceinline__ uint32_t bytes4(const BlobView& v, uint32_t at)
{
    const uint32_t w = at >> 2, sh = (at & 3u) * 8u;
    const uint32_t lo = pay_word(v, w);
    if (sh == 0u) return lo;
    const uint32_t hi = pay_word(v, w + 1u);
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> sh);
}

__device__ __forceinline__ float mixf(float a, float b, float t)
{
    const float p = a * (1.0f - t);
    const float q = b * t;
    return p + q;
}

__device__ __forceinline__ uint32_t key_frame(const BlobView& v, uint32_t keys, uint32_t codec, uint32_t k)
{
    return codec == 1u ? bytes4(v, keys + 8u * k) & 0xFFFFu : bytes4(v, keys + 3u * k) & 0xFFu;
}

__device__ __forceinline__ float key_value(const BlobView& v, uint32_t keys, uint32_t codec, uint32_t k)
{
    if (codec == 1u) return __uint_as_float(bytes4(v, keys + 8u * k + 4u));
    return half_to_float((bytes4(v, keys + 3u * k) >> 8) & 0xFFFFu);
}

// decode_temporal_control_value (vbt_cycles_sampler.h:165-212; the patch's
// copy, native_cycles_vbt.patch:639-685, is the same)
__device__ __forceinline__ float control_value(const BlobView& v, uint32_t stream, uint32_t codec, uint32_t nc, uint32_t nkeys,
                               uint32_t ci)
{
    const uint32_t descBytes = nc * 4u;
    const uint32_t keys = stream + descBytes + ((nc * 8u + 3u) & ~3u);
    const uint32_t desc = bytes4(v, stream + ci * 4u);
    const uint32_t start = desc & 0xFFFFu, count = (desc >> 16) & 0xFFu;
    if (count == 0u || start + count > nkeys) return 0.0f;
    const int F = (int)v.frames;
    const int last = F > 0 ? F - 1 : 0;
    const uint32_t t = (uint32_t)(v.frameIndex < 0 ? 0 : (v.frameIndex > last ? last : v.frameIndex));
    const uint32_t bin = min(7u, t * 8u / max(v.frames, 1u));
    const uint32_t hint = bytes4(v, stream + descBytes + ci * 8u + bin);  // bins[bin], bins[bin + 1]
    uint32_t lo = min(hint & 0xFFu, count - 1u);
    uint32_t f0 = key_frame(v, keys, codec, start + lo);
    if (f0 > t) {
        lo = 0u;
        f0 = key_frame(v, keys, codec, start);
    }
    if (f0 >= t || lo + 1u >= count) return key_value(v, keys, codec, start + lo);
    uint32_t hi = bin < 7u ? min(count - 1u, ((hint >> 8) & 0xFFu) + 1u) : count - 1u;
    hi = max(hi, lo + 1u);
    uint32_t f1 = key_frame(v, keys, codec, start + hi);
    if (f1 < t) {
        hi = count - 1u;
        f1 = key_frame(v, keys, codec, start + hi);
    }
    if (f1 <= t) return key_value(v, keys, codec, start + hi);
    while (hi > lo + 1u) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        const uint32_t fm = key_frame(v, keys, codec, start + mid);
        if (fm < t) {
            lo = mid;
            f0 = fm;
        } else {
            hi = mid;
            f1 = fm;
        }
    }
    const float next = key_value(v, keys, codec, start + hi);
    if (f1 == t) return next;
    const float prev = key_value(v, keys, codec, start + lo);
    const float alpha = (float)(uint32_t)(t - f0) / (float)(uint32_t)(f1 - f0);
    return mixf(prev, next, alpha);
}

__device__ __forceinline__ float ctrl_pos(uint32_t i, uint32_t res)
{
    return res <= 1u ? 0.0f : (float)i * (7.0f / (float)(res - 1u));
}

// locate_coord (vbt_cycles_sampler.h:219-243); the kernel form
// (native_cycles_vbt.patch:694-719) sends resolution <= 1 to (0, 0, 0)
template <bool KF>
__device__ __forceinline__ void locate(float p, uint32_t extent, uint32_t res, uint32_t& i0, uint32_t& i1, float& w)
{
    if (KF && res <= 1u) {
        i0 = i1 = 0u;
        w = 0.0f;
        return;
    }
    const float maxP = (float)(extent > 0u ? extent - 1u : 0u);
    const float c = p < 0.0f ? 0.0f : (p > maxP ? maxP : p);
    uint32_t hi = 1u;
    while (hi < res - 1u && ctrl_pos(hi, res) < c) ++hi;
    const uint32_t lo = hi > 0u ? hi - 1u : 0u;
    const float c0 = ctrl_pos(lo, res), c1 = ctrl_pos(hi, res);
    i0 = lo;
    i1 = hi;
    w = c1 > c0 ? (c - c0) / (c1 - c0) : 0.0f;
}

// Combines the values of 8 corners held by lanes c = 0..7 at lane stride S
// (corner bit 0 = x, 1 = y, 2 = z) in the reference's order:
// mix(mix(mix(v000, v100), mix(v010, v110)), mix(mix(v001, v101), mix(v011, v111))).
// Every lane of the group ends with the result.
template <uint32_t S>
__device__ __forceinline__ float corner_mix(float val, uint32_t c, float wx, float wy, float wz)
{
    float o = __shfl_xor(val, (int)S);
    float m = (c & 1u) ? mixf(o, val, wx) : mixf(val, o, wx);
    o = __shfl_xor(m, (int)(2u * S));
    m = (c & 2u) ? mixf(o, m, wy) : mixf(m, o, wy);
    o = __shfl_xor(m, (int)(4u * S));
    return (c & 4u) ? mixf(o, m, wz) : mixf(m, o, wz);
}

__device__ __forceinline__ float serial_mix(const float* c, float wx, float wy, float wz)
{
    const float c00 = mixf(c[0], c[1], wx), c10 = mixf(c[2], c[3], wx);
    const float c01 = mixf(c[4], c[5], wx), c11 = mixf(c[6], c[7], wx);
    return mixf(mixf(c00, c10, wy), mixf(c01, c11, wy), wz);
}

// sample_control_stream (vbt_cycles_sampler.h:245-282).  G = 8: lane c of an
// aligned group of 8 decodes lattice corner c; G = 1: one lane decodes all 8.
template <bool KF, uint32_t G>
__device__ __forceinline__ float stream_sample(const BlobView& v, uint32_t stream, uint32_t codec, uint32_t res, uint32_t nkeys,
                               uint32_t ex, uint32_t ey, uint32_t ez, float lx, float ly, float lz, uint32_t c)
{
    if (res == 0u) return 0.0f;
    uint32_t x0, x1, y0, y1, z0, z1;
    float wx, wy, wz;
    locate<KF>(lx, ex, res, x0, x1, wx);
    locate<KF>(ly, ey, res, y0, y1, wy);
    locate<KF>(lz, ez, res, z0, z1, wz);
    const uint32_t nc = res * res * res;
    if constexpr (G == 8u) {
        const uint32_t x = (c & 1u) ? x1 : x0, y = (c & 2u) ? y1 : y0, z = (c & 4u) ? z1 : z0;
        const float val = control_value(v, stream, codec, nc, nkeys, (z * res + y) * res + x);
        return corner_mix<1u>(val, c, wx, wy, wz);
    } else {
        float cv[8];
#pragma unroll
        for (uint32_t k = 0; k < 8u; ++k) {
            const uint32_t x = (k & 1u) ? x1 : x0, y = (k & 2u) ? y1 : y0, z = (k & 4u) ? z1 : z0;
            cv[k] = control_value(v, stream, codec, nc, nkeys, (z * res + y) * res + x);
        }
        return serial_mix(cv, wx, wy, wz);
    }
}

// sample_density_index (vbt_cycles_sampler.h:284-368); with KF the kernel
// form's vbt_sample_density_leaf_index (native_cycles_vbt.patch:788-874):
// Mode2 leaves give 0 and leaf_count_z comes from the header.
template <bool KF, uint32_t G>
__device__ __forceinline__ float leaf_sample(const BlobView& v, float x, float y, float z, uint32_t c)
{
    if (x < 0.0f || y < 0.0f || z < 0.0f || x > (float)(v.dimX - 1u) || y > (float)(v.dimY - 1u) ||
        z > (float)(v.dimZ - 1u))
        return 0.0f;
    const uint32_t xi = min(v.dimX - 1u, (uint32_t)x), yi = min(v.dimY - 1u, (uint32_t)y),
                   zi = min(v.dimZ - 1u, (uint32_t)z);
    const uint32_t S = v.leafSize;
    const uint32_t lcz = KF ? v.lcz : (v.dimZ + S - 1u) / S;
    const uint32_t bx = min(v.lcx - 1u, xi / S), by = min(v.lcy - 1u, yi / S), bz = min(lcz - 1u, zi / S);
    const uint32_t leaf = (bz * v.lcy + by) * v.lcx + bx;
    const uint32_t wb = leaf < v.offWords ? v.off[leaf] : 0u;
    const uint32_t we = leaf + 1u < v.offWords ? v.off[leaf + 1u] : 0u;
    if (we <= wb) return 0.0f;
    const uint32_t base = wb * 4u;
    const uint32_t hdr = pay_word(v, wb);
    const uint32_t mode = hdr & 3u;
    if (mode == 0u || (KF && mode == 2u)) return 0.0f;
    const uint32_t cc = (hdr >> 2) & 3u, fc = (hdr >> 4) & 3u, cr = (hdr >> 6) & 7u, fr = (hdr >> 9) & 7u;
    const uint32_t keyCounts = pay_word(v, wb + 1u);
    const uint32_t ck = keyCounts & 0xFFFFu, fk = keyCounts >> 16;
    const uint32_t cn = cr * cr * cr;
    const uint32_t cStream = base + 8u;
    const uint32_t fStream =
        cStream + cn * 4u + ((cn * 8u + 3u) & ~3u) + ((ck * (cc == 1u ? 8u : 3u) + 3u) & ~3u);
    const uint32_t ox = bx * S, oy = by * S, oz = bz * S;
    const uint32_t ex = min(S, v.dimX - ox), ey = min(S, v.dimY - oy), ez = min(S, v.dimZ - oz);
    const float lx = x - (float)ox, ly = y - (float)oy, lz = z - (float)oz;
    float value = stream_sample<KF, G>(v, cStream, cc, cr, ck, ex, ey, ez, lx, ly, lz, c);
    if (mode == 3u && fr > 0u) value += stream_sample<KF, G>(v, fStream, fc, fr, fk, ex, ey, ez, lx, ly, lz, c);
    return value;
}

// Standalone form: G lanes per sample.
template <uint32_t G>
__global__ void __launch_bounds__(256) k_cycles_standalone(const uint32_t* blob, uint64_t bytes, int world,
                                                           const float* xyz, uint32_t n, float* out)
{
    const uint64_t gt = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t i = gt / G;
    if (i >= n) return;
    const uint32_t c = (uint32_t)(gt % G);
    BlobView v;
    if (!open_view(blob, bytes, v)) {
        if (c == 0u) out[i] = 0.0f;
        return;
    }
    float x = xyz[3u * i], y = xyz[3u * i + 1u], z = xyz[3u * i + 2u];
    if (world) {
        x -= v.bx;
        y -= v.by;
        z -= v.bz;
    }
    const float r = leaf_sample<false, G>(v, x, y, z, c);
    if (c == 0u) out[i] = r;
}

// Kernel form (native_cycles_vbt.patch:876-913): the eight voxel corners
// around the point.  G = 8: 64 lanes per sample, lane = 8 * voxel corner +
// lattice corner; G = 1: one lane per sample.
template <uint32_t G>
__global__ void __launch_bounds__(256) k_cycles_kernel(const uint32_t* blob, uint64_t bytes, int world,
                                                       const float* xyz, uint32_t n, float* out)
{
    constexpr uint32_t LPS = G * 8u == 64u ? 64u : 1u;  // lanes per sample
    const uint64_t gt = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t i = gt / LPS;
    if (i >= n) return;
    const uint32_t lane = (uint32_t)(gt % LPS);
    BlobView v;
    if (!open_view(blob, bytes, v)) {
        if (lane == 0u) out[i] = 0.0f;
        return;
    }
    float x = xyz[3u * i], y = xyz[3u * i + 1u], z = xyz[3u * i + 2u];
    if (world) {
        x -= v.bx;
        y -= v.by;
        z -= v.bz;
    }
    if (x < 0.0f || y < 0.0f || z < 0.0f || x > (float)(v.dimX - 1u) || y > (float)(v.dimY - 1u) ||
        z > (float)(v.dimZ - 1u)) {
        if (lane == 0u) out[i] = 0.0f;
        return;
    }
    const uint32_t x0 = min(v.dimX - 1u, (uint32_t)x), y0 = min(v.dimY - 1u, (uint32_t)y),
                   z0 = min(v.dimZ - 1u, (uint32_t)z);
    const uint32_t x1 = min(v.dimX - 1u, x0 + 1u), y1 = min(v.dimY - 1u, y0 + 1u), z1 = min(v.dimZ - 1u, z0 + 1u);
    const float tx = x1 > x0 ? x - (float)x0 : 0.0f;
    const float ty = y1 > y0 ? y - (float)y0 : 0.0f;
    const float tz = z1 > z0 ? z - (float)z0 : 0.0f;
    float r;
    if constexpr (G == 8u) {
        const uint32_t s = lane >> 3, c = lane & 7u;
        const float val = leaf_sample<true, 8u>(v, (float)((s & 1u) ? x1 : x0), (float)((s & 2u) ? y1 : y0),
                                                (float)((s & 4u) ? z1 : z0), c);
        r = corner_mix<8u>(val, s, tx, ty, tz);
    } else {
        float cv[8];
        for (uint32_t k = 0; k < 8u; ++k)
            cv[k] = leaf_sample<true, 1u>(v, (float)((k & 1u) ? x1 : x0), (float)((k & 2u) ? y1 : y0),
                                          (float)((k & 4u) ? z1 : z0), 0u);
        r = serial_mix(cv, tx, ty, tz);
    }
    if (lane == 0u) out[i] = r;
}

int launch_check(const char* what)
{
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? VBT_OK : vbt_fail(VBT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" {

int vbt_cycles_blob_bytes(vbt_asset a, uint64_t* bytes)
{
    if (!a || !bytes) return vbt_fail(VBT_ERR_INVALID, "vbt_cycles_blob_bytes: null argument");
    const vbt_file_header* h = vbt_internal_asset_header(a);
    *bytes = BLOB_HEADER_BYTES + 4ull * (h->leafCount + 1ull) + vbt_internal_asset_dev(a).payloadBytes;
    return VBT_OK;
}

int vbt_cycles_pack_blob(vbt_asset a, int frameIndex, void* d_blob, uint64_t capacity, void* stream)
{
    if (!a || !d_blob) return vbt_fail(VBT_ERR_INVALID, "vbt_cycles_pack_blob: null argument");
    if ((uintptr_t)d_blob % 4u) return vbt_fail(VBT_ERR_INVALID, "vbt_cycles_pack_blob: blob must be 4-byte aligned");
    const vbt_file_header* h = vbt_internal_asset_header(a);
    // pack_blob's own checks (vbt_cycles_blob.cpp:24-50); an open asset has
    // already passed the loader's (vbt_cycles_loader.cpp:36-221)
    if (!h->width || !h->height || !h->depth || !h->frames || !h->leafCount)
        return vbt_fail(VBT_ERR_FORMAT, "Invalid empty VBT dimensions or frame count.");
    const uint32_t S = h->leafSize ? h->leafSize : 8u;
    const uint64_t lcx = (h->width + S - 1ull) / S, lcy = (h->height + S - 1ull) / S, lcz = (h->depth + S - 1ull) / S;
    if (lcx * lcy * lcz != h->leafCount) return vbt_fail(VBT_ERR_FORMAT, "Computed leaf count does not match VBT header.");
    const AssetDev ad = vbt_internal_asset_dev(a);
    const uint64_t payBytes = ad.payloadBytes;
    const uint64_t offAt = BLOB_HEADER_BYTES, payAt = offAt + 4ull * (h->leafCount + 1ull);
    if (payAt > 0xFFFFFFFFull) return vbt_fail(VBT_ERR_FORMAT, "VBT blob section offset exceeds the 32-bit header range.");
    if (capacity < payAt + payBytes) return vbt_fail(VBT_ERR_INVALID, "vbt_cycles_pack_blob: blob buffer too small");
    const int last = (int)h->frames - 1;
    BlobHeaderWords hw;
    const uint32_t w[16] = {BLOB_MAGIC, 1u, h->width, h->height, h->depth, h->frames, S, (uint32_t)lcx, (uint32_t)lcy,
                            (uint32_t)lcz, h->leafCount, (uint32_t)(frameIndex < 0 ? 0 : (frameIndex > last ? last : frameIndex)),
                            h->leafCount + 1u, (uint32_t)(payBytes / 4u), (uint32_t)offAt, (uint32_t)payAt};
    for (int k = 0; k < 16; ++k) hw.w[k] = w[k];
    if (hipSetDevice(vbt_internal_asset_device(a)) != hipSuccess) return vbt_fail(VBT_ERR_DEVICE, "hipSetDevice");
    hipStream_t s = (hipStream_t)stream;
    uint8_t* dst = (uint8_t*)d_blob;
    k_blob_header<<<1, 32, 0, s>>>((uint32_t*)dst, hw);
    int rc = launch_check("vbt_cycles_pack_blob");
    if (rc) return rc;
    if (hipMemcpyAsync(dst + offAt, ad.offsets, 4ull * (h->leafCount + 1ull), hipMemcpyDeviceToDevice, s) != hipSuccess ||
        (payBytes && hipMemcpyAsync(dst + payAt, ad.payload, payBytes, hipMemcpyDeviceToDevice, s) != hipSuccess))
        return vbt_fail(VBT_ERR_DEVICE, "vbt_cycles_pack_blob: device copy failed");
    return VBT_OK;
}

int vbt_cycles_sample_density(const void* d_blob, uint64_t blobBytes, int flags, const float* d_xyz, uint32_t n,
                              float* d_out, void* stream)
{
    if (n == 0u) return VBT_OK;
    if (!d_blob || !d_xyz || !d_out) return vbt_fail(VBT_ERR_INVALID, "vbt_cycles_sample_density: null buffer");
    if ((uintptr_t)d_blob % 4u) return vbt_fail(VBT_ERR_INVALID, "vbt_cycles_sample_density: blob must be 4-byte aligned");
    if (flags & ~(VBT_CYCLES_KERNEL | VBT_CYCLES_WORLD | VBT_CYCLES_PLAIN))
        return vbt_fail(VBT_ERR_INVALID, "vbt_cycles_sample_density: unknown flag");
    const bool kf = (flags & VBT_CYCLES_KERNEL) != 0, plain = (flags & VBT_CYCLES_PLAIN) != 0;
    const int world = (flags & VBT_CYCLES_WORLD) != 0;
    const uint64_t lanes = (uint64_t)n * (plain ? 1u : (kf ? 64u : 8u));
    if (lanes > 0xFFFFFF00ull) return vbt_fail(VBT_ERR_INVALID, "vbt_cycles_sample_density: too many samples for one launch");
    const uint32_t blocks = (uint32_t)((lanes + 255u) / 256u);
    hipStream_t s = (hipStream_t)stream;
    const uint32_t* b = (const uint32_t*)d_blob;
    if (kf) {
        if (plain) k_cycles_kernel<1u><<<blocks, 256, 0, s>>>(b, blobBytes, world, d_xyz, n, d_out);
        else k_cycles_kernel<8u><<<blocks, 256, 0, s>>>(b, blobBytes, world, d_xyz, n, d_out);
    } else {
        if (plain) k_cycles_standalone<1u><<<blocks, 256, 0, s>>>(b, blobBytes, world, d_xyz, n, d_out);
        else k_cycles_standalone<8u><<<blocks, 256, 0, s>>>(b, blobBytes, world, d_xyz, n, d_out);
    }
    return launch_check("vbt_cycles_sample_density");
}

}  // extern "C"
