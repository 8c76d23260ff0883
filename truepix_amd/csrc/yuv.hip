// yuv.hip - the frame's macroblock layout on the device: YUV 4:2:0 planes to
// the prover's per-copy input rows and back.
//
// Packing follows VideoEdition/Frame_Block_Extraction.py:67-94
// (process_macroblocks) and :220-240 (process_video_and_save_to_txt, the
// input.txt line per macroblock): macroblocks in raster order, width / 16 per
// row and height / 16 rows (a partial macroblock at the right or bottom edge
// is dropped), each copy the 256 bytes of its 16 x 16 Y block row by row, then
// the 64 of its 8 x 8 Cb block, then the 64 of its 8 x 8 Cr block.  The crop
// edit appends one byte per copy, the macroblock's crop bit
// (VideoEdition/Crop.py:25-37, written after each line at :79-85).
// Unpacking is VideoEdition/Encode.py:19-42 (reconstruct_frame): every
// macroblock's bytes back into freshly zeroed planes, so pixels outside the
// macroblock area come out 0.
//
// Both directions are byte shuffles bound by HBM: one thread per output byte,
// consecutive threads writing consecutive bytes; the reads run along plane
// rows (16 or 8 bytes) or along a copy's row.
#include "runtime.hpp"

#include <hip/hip_runtime.h>
#include <stdint.h>

using namespace tp;

namespace {

constexpr int kMb = 16;        // macroblock side (Y)
constexpr int kCopy = 384;     // 256 Y + 64 Cb + 64 Cr

__global__ void __launch_bounds__(256) k_yuv_pack(const uint8_t* __restrict__ y,
                                                  const uint8_t* __restrict__ cb,
                                                  const uint8_t* __restrict__ cr, int mbx,
                                                  long long y_stride, long long c_stride,
                                                  const uint8_t* __restrict__ extra, int copy_bytes,
                                                  long long total, uint8_t* __restrict__ out)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const long long c = i / copy_bytes;
        const int k = (int)(i - c * copy_bytes);
        const long long my = c / mbx, mx = c - (c / mbx) * mbx;
        uint8_t v;
        if (k < 256) {
            v = y[(my * kMb + (k >> 4)) * y_stride + mx * kMb + (k & 15)];
        } else if (k < 320) {
            const int j = k - 256;
            v = cb[(my * 8 + (j >> 3)) * c_stride + mx * 8 + (j & 7)];
        } else if (k < kCopy) {
            const int j = k - 320;
            v = cr[(my * 8 + (j >> 3)) * c_stride + mx * 8 + (j & 7)];
        } else {
            v = extra[c];
        }
        out[i] = v;
    }
}

// One thread per plane byte of the three planes, Y first.
__global__ void __launch_bounds__(256) k_yuv_unpack(const uint8_t* __restrict__ mb,
                                                    long long copy_stride, int mbx, int mby,
                                                    int width, int height, long long y_stride,
                                                    long long c_stride, uint8_t* __restrict__ y,
                                                    uint8_t* __restrict__ cb,
                                                    uint8_t* __restrict__ cr)
{
    const long long ny = (long long)width * height;
    const int cw = width / 2, ch = height / 2;
    const long long nc = (long long)cw * ch;
    const long long total = ny + 2 * nc;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        if (i < ny) {
            const int row = (int)(i / width), col = (int)(i - (long long)row * width);
            uint8_t v = 0;
            if (row < mby * kMb && col < mbx * kMb) {
                const long long c = (long long)(row >> 4) * mbx + (col >> 4);
                v = mb[c * copy_stride + (row & 15) * 16 + (col & 15)];
            }
            y[(long long)row * y_stride + col] = v;
        } else {
            const long long j = i - ny;
            const int plane = j >= nc;   // 0 Cb, 1 Cr
            const long long jj = j - plane * nc;
            const int row = (int)(jj / cw), col = (int)(jj - (long long)row * cw);
            uint8_t v = 0;
            if (row < mby * 8 && col < mbx * 8) {
                const long long c = (long long)(row >> 3) * mbx + (col >> 3);
                v = mb[c * copy_stride + 256 + plane * 64 + (row & 7) * 8 + (col & 7)];
            }
            (plane ? cr : cb)[(long long)row * c_stride + col] = v;
        }
    }
}

unsigned grid_bytes(long long n)
{
    long long b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b > 0 ? b : 1);
}

bool check_frame(const char* who, int width, int height, long long y_stride, long long c_stride)
{
    if (width < kMb || height < kMb || (width & 1) || (height & 1))
        return fail("%s: frame %dx%d (needs even sides of at least 16)", who, width, height);
    if (y_stride < width || c_stride < width / 2)
        return fail("%s: strides %lld/%lld below the plane widths", who, y_stride, c_stride);
    return true;
}

// Staging for the host-buffer entry points: planes and copies on the device.
struct YuvStage {
    DevBuf planes, copies, extra;
};
YuvStage g_yuv;

}  // namespace

extern "C" {

// Packs a frame's planes into per-copy input rows on the device.
//   y (height rows of y_stride bytes), cb and cr (height/2 rows of c_stride
//   bytes): device pointers; extra: one byte per copy appended as byte 384
//   when copy_bytes is 385 (the crop bit), else null; out: copies x
//   copy_bytes bytes, device.  stream: hipStream_t or NULL.
// Returns the number of copies ((width/16) * (height/16)), or -1.
long long tp_yuv_pack(const uint8_t* y, const uint8_t* cb, const uint8_t* cr, int width,
                      int height, long long y_stride, long long c_stride, const uint8_t* extra,
                      int copy_bytes, uint8_t* out, void* stream)
{
    clear_error();
    if (!check_frame("tp_yuv_pack", width, height, y_stride, c_stride)) return -1;
    if (copy_bytes != kCopy && copy_bytes != kCopy + 1)
        return fail("tp_yuv_pack: copy_bytes=%d (384, or 385 with a per-copy byte)", copy_bytes), -1;
    if ((copy_bytes == kCopy + 1) != (extra != nullptr))
        return fail("tp_yuv_pack: a per-copy byte goes with copy_bytes=385 and only then"), -1;
    if (!y || !cb || !cr || !out) return fail("tp_yuv_pack: null plane or output"), -1;
    if (!device_ready()) return -1;
    const int mbx = width / kMb, mby = height / kMb;
    const long long copies = (long long)mbx * mby;
    const long long total = copies * copy_bytes;
    hipStream_t st = stream ? (hipStream_t)stream : tp::stream();
    k_yuv_pack<<<grid_bytes(total), 256, 0, st>>>(y, cb, cr, mbx, y_stride, c_stride, extra,
                                                  copy_bytes, total, out);
    TP_HIP_OK(hipGetLastError());
    return copies;
}

// Writes the planes of a frame from its per-copy rows on the device (zeros
// outside the macroblock area).  mb: copies x copy_stride bytes (the first
// 384 of each row used), device; y, cb, cr as for tp_yuv_pack.  0 / -1.
int tp_yuv_unpack(const uint8_t* mb, long long copy_stride, int width, int height,
                  long long y_stride, long long c_stride, uint8_t* y, uint8_t* cb, uint8_t* cr,
                  void* stream)
{
    clear_error();
    if (!check_frame("tp_yuv_unpack", width, height, y_stride, c_stride)) return -1;
    if (copy_stride < kCopy) return fail("tp_yuv_unpack: copy_stride=%lld < 384", copy_stride), -1;
    if (!mb || !y || !cb || !cr) return fail("tp_yuv_unpack: null pointer"), -1;
    if (!device_ready()) return -1;
    const int mbx = width / kMb, mby = height / kMb;
    const long long total = (long long)width * height + 2LL * (width / 2) * (height / 2);
    hipStream_t st = stream ? (hipStream_t)stream : tp::stream();
    k_yuv_unpack<<<grid_bytes(total), 256, 0, st>>>(mb, copy_stride, mbx, mby, width, height,
                                                    y_stride, c_stride, y, cb, cr);
    TP_HIP_OK(hipGetLastError());
    return 0;
}

// tp_yuv_pack from host buffers (tightly packed planes: y_stride = width,
// c_stride = width/2) to a host buffer of copies x copy_bytes, through the
// device; synchronous.  Returns the number of copies, or -1.
long long tp_yuv_pack_host(const uint8_t* y, const uint8_t* cb, const uint8_t* cr, int width,
                           int height, const uint8_t* extra, int copy_bytes, uint8_t* out)
{
    clear_error();
    if (!check_frame("tp_yuv_pack_host", width, height, width, width / 2)) return -1;
    if (!y || !cb || !cr || !out) return fail("tp_yuv_pack_host: null pointer"), -1;
    if (!device_ready()) return -1;
    const size_t ny = (size_t)width * height, nc = (size_t)(width / 2) * (height / 2);
    const long long copies = (long long)(width / kMb) * (height / kMb);
    YuvStage& S = g_yuv;
    if (!S.planes.ensure(ny + 2 * nc) || !S.copies.ensure((size_t)copies * 385) ||
        !S.extra.ensure((size_t)copies))
        return -1;
    hipStream_t st = tp::stream();
    uint8_t* dp = S.planes.as<uint8_t>();
    TP_HIP_OK(hipMemcpyAsync(dp, y, ny, hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipMemcpyAsync(dp + ny, cb, nc, hipMemcpyHostToDevice, st));
    TP_HIP_OK(hipMemcpyAsync(dp + ny + nc, cr, nc, hipMemcpyHostToDevice, st));
    const uint8_t* dextra = nullptr;
    if (extra) {
        TP_HIP_OK(hipMemcpyAsync(S.extra.as<uint8_t>(), extra, (size_t)copies, hipMemcpyHostToDevice, st));
        dextra = S.extra.as<uint8_t>();
    }
    const long long n = tp_yuv_pack(dp, dp + ny, dp + ny + nc, width, height, width, width / 2,
                                    dextra, copy_bytes, S.copies.as<uint8_t>(), st);
    if (n < 0) return -1;
    TP_HIP_OK(hipMemcpyAsync(out, S.copies.as<uint8_t>(), (size_t)n * copy_bytes,
                             hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    return n;
}

// tp_yuv_unpack from a host buffer to tightly packed host planes, through
// the device; synchronous.  0 / -1.
int tp_yuv_unpack_host(const uint8_t* mb, long long copy_stride, int width, int height, uint8_t* y,
                       uint8_t* cb, uint8_t* cr)
{
    clear_error();
    if (!check_frame("tp_yuv_unpack_host", width, height, width, width / 2)) return -1;
    if (copy_stride < kCopy) return fail("tp_yuv_unpack_host: copy_stride=%lld < 384", copy_stride), -1;
    if (!mb || !y || !cb || !cr) return fail("tp_yuv_unpack_host: null pointer"), -1;
    if (!device_ready()) return -1;
    const size_t ny = (size_t)width * height, nc = (size_t)(width / 2) * (height / 2);
    const long long copies = (long long)(width / kMb) * (height / kMb);
    YuvStage& S = g_yuv;
    if (!S.planes.ensure(ny + 2 * nc) || !S.copies.ensure((size_t)copies * copy_stride)) return -1;
    hipStream_t st = tp::stream();
    uint8_t* dmb = S.copies.as<uint8_t>();
    uint8_t* dp = S.planes.as<uint8_t>();
    TP_HIP_OK(hipMemcpyAsync(dmb, mb, (size_t)copies * copy_stride, hipMemcpyHostToDevice, st));
    if (tp_yuv_unpack(dmb, copy_stride, width, height, width, width / 2, dp, dp + ny, dp + ny + nc,
                      st) != 0)
        return -1;
    TP_HIP_OK(hipMemcpyAsync(y, dp, ny, hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipMemcpyAsync(cb, dp + ny, nc, hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipMemcpyAsync(cr, dp + ny + nc, nc, hipMemcpyDeviceToHost, st));
    TP_HIP_OK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
