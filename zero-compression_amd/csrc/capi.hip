// The C ABI of include/zero_hip.h: device context, memory, streams, and the launch sequences of the encoder and
// decoder (the roles of src/gpu.h and of GpuCodec::compress / decompress in src/gpu_codec.h:153-315).
// One translation unit holds every kernel, so libzero_hip.so is a single code object for gfx950.
#include "crc.hip"
#include "decode.hip"
#include "encode.hip"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/zero_hip.h"

using namespace zk;

namespace {

// a device buffer that grows on demand
struct DBuf {
  void* p = nullptr;
  size_t n = 0;
};

struct Stage {
  const char* name;
  hipEvent_t ev;
};

}  // namespace

// The decoder's scratch for zgpu_decode_blocks_noalloc: a fixed array in the code object, so that a call allocates
// nothing. Each GPU the library runs on holds its own copy (ZGPU_NOALLOC_SCRATCH_BYTES) until the process ends. In
// host code `scratch` names a host-side stand-in: its device address on each GPU comes from hipGetSymbolAddress
// (noalloc_scratch_of), and only that address is ever handed to kernels.
static __device__ __align__(256) uint8_t scratch[ZGPU_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {};  // device address of `scratch` on each GPU, once looked up
// the address of `scratch` on device dev (current), looked up once per device; nullptr on failure
hipError_t noalloc_scratch_of(int dev, void** out) {
  *out = nullptr;
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return hipSuccess;
}
}  // namespace

// one of the two streams of the host-buffer forms, with what a chunk of a batch needs on it
struct Lane {
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;      // compress: the chunk's lengths and payload total are on the host
  uint64_t* h_total = nullptr;  // (pinned) the chunk's payload bytes
  DBuf work, slab, pay, meta;   // scratch; compress: slabs, packed payloads; block table, lengths, CRCs, offsets
};

struct zgpu {
  int dev = 0;
  void* noalloc = nullptr;  // this device's `scratch` (its device address)
  Lane lanes[2];
  DBuf hin, hout;           // host-buffer forms: the whole batch's input (compress, decompress) and output
  hipStream_t stream = nullptr;
  std::string err;
  zgpu_props_t props{};
  uint32_t* xpow = nullptr;  // x^(8 m 256^j) mod P, 3 x 256 (CRC combine)
  DBuf enc, dec, aux;  // encoder and decoder scratch of the device-resident forms; staging of the small host forms
  DBuf rng;            // byte-range CRC tables
  DBuf stage;          // (pinned host memory) the host-buffer forms' block tables and per-block results
  // the last zgpu_compress input [0] (in hin) and zgpu_decompress output [1] (in hout); length 0: none
  const uint8_t* last_data[2] = {nullptr, nullptr};
  uint64_t last_len[2] = {0, 0};
  bool profile = false;
  std::vector<Stage> stages;
  std::vector<hipEvent_t> pool;
  size_t used_events = 0;
};

static int fail(zgpu_t* g, int code, const std::string& msg) {
  if (g) g->err = msg;
  return code;
}
static int hip_fail(zgpu_t* g, hipError_t e, const char* what) {
  const int code = e == hipErrorOutOfMemory ? ZGPU_ERR_NOMEM : ZGPU_ERR_DEVICE;
  // device loss and launch failures read as the reference's "the GPU was reset" (src/gpu.h:20-27)
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorLaunchFailure || e == hipErrorIllegalAddress || e == hipErrorECCNotCorrectable)
    m += " (the GPU was reset)";
  return fail(g, code, m);
}
#define HIPCHK(g, call)                                  \
  do {                                                   \
    const hipError_t e_ = (call);                        \
    if (e_ != hipSuccess) return hip_fail(g, e_, #call); \
  } while (0)

// pinned host memory for the copies of a pipelined call: an asynchronous copy to or from pageable memory makes the
// host wait until the copy is done, so the next chunk would not be queued until this one had finished
static int ensure_pinned(zgpu_t* g, DBuf& b, size_t bytes) {
  if (b.n >= bytes) return ZGPU_OK;
  if (b.p) HIPCHK(g, hipHostFree(b.p));
  b.p = nullptr;
  b.n = 0;
  HIPCHK(g, hipHostMalloc(&b.p, bytes, hipHostMallocDefault));
  b.n = bytes;
  return ZGPU_OK;
}
static int ensure(zgpu_t* g, DBuf& b, size_t bytes) {
  if (b.n >= bytes) return ZGPU_OK;
  if (b.p) HIPCHK(g, hipFree(b.p));
  b.p = nullptr;
  b.n = 0;
  HIPCHK(g, hipMalloc(&b.p, bytes));
  b.n = bytes;
  return ZGPU_OK;
}

static void stage_begin(zgpu_t* g, hipStream_t s) {
  g->stages.clear();
  g->used_events = 0;
  if (!g->profile) return;
  if (g->pool.empty()) {
    for (int i = 0; i < 32; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      g->pool.push_back(e);
    }
  }
  g->stages.push_back({"start", g->pool[g->used_events++]});
  (void)hipEventRecord(g->stages.back().ev, s);
}
static void stage(zgpu_t* g, hipStream_t s, const char* name) {
  if (!g->profile || g->used_events >= g->pool.size()) return;
  g->stages.push_back({name, g->pool[g->used_events++]});
  (void)hipEventRecord(g->stages.back().ev, s);
}

// ---------------------------------------------------------------- small kernels
__global__ void k_fill(uint8_t* p, uint8_t v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = i; k < n; k += stride) p[k] = v;
}

// zeros over n16 16-byte words from p (16-byte aligned)
__global__ __launch_bounds__(256) void k_zero16(uint4* __restrict__ p, size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) p[i] = uint4{0u, 0u, 0u, 0u};
}
// a plain copy, 16 bytes per lane per step (both pointers 16-byte aligned, n a multiple of 16; the host does the rest)
__global__ __launch_bounds__(256) void k_copy(uint4* __restrict__ d, const uint4* __restrict__ s, size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += stride) d[k] = s[k];
}

// ---------------------------------------------------------------- scratch layouts
namespace {
inline size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

struct EncLayout {
  size_t srt, mtch, bstart, qmid, qlarge, qcnt, mark, lit, LB, TM, TD, tok, sym, grp, ext, offw, cnt, lanes, freq,
      lanelen, sinfo, total;
  uint32_t qcap;
};
// The launch shape of an encoder call: slabs of 2^blog bytes (every block fits its slab), buckets of the top hbits
// hash bits, keys of keybits bits, `walkers` parse segments per block. The bytes depend on keybits alone (the
// archive's block size decides it); blog, hbits and walkers change only the work.
struct EncShape {
  uint32_t blog, hbits, keybits, walkers;
};
// the parse's segments per block: one per PARSE_SPAN positions but at least PARSE_WALKERS_MIN, each at least 32
// positions, and the parse's LDS (3 words and a window per walker) within 160 KiB.
// Fewer, longer segments cost more serial walking; more, shorter ones cost more joins (a long match crosses many).
uint32_t walkers_of(uint32_t blog) {
  uint32_t w = std::min<uint32_t>(std::max<uint32_t>(PARSE_WALKERS_MIN, (1u << blog) / PARSE_SPAN), (1u << blog) / 32);
  while (w > 1 && parse_lds_words(blog, w) * 4 > 160 * 1024) w >>= 1;
  return w;
}
EncShape default_shape(uint32_t blog) { return {blog, hbits_of(blog), hbits_of(blog) + KEY_TAG_BITS, walkers_of(blog)}; }
bool shape_ok(const EncShape& S) {
  return S.blog >= 10 && S.blog <= 20 && S.hbits >= 4 && S.hbits <= 15 && S.hbits <= S.keybits &&
         S.keybits <= 27 && S.keybits - S.hbits + S.blog <= 32 && S.walkers >= 1 && S.walkers <= 1024 &&
         (S.walkers & (S.walkers - 1)) == 0 && ((1u << S.blog) / S.walkers) >= 32;
}

EncLayout enc_layout(uint32_t nb, uint32_t blog, uint32_t hbits) {
  const Geo G = geo_of(blog);
  const size_t N = (size_t)nb << blog;
  const size_t HB = (size_t)1 << hbits;
  EncLayout L;
  size_t o = 0;
  L.qcap = (uint32_t)((size_t)nb * (G.B / 17u + 1u));
  L.srt = o; o += al(N * 4);                // bucketed positions, then sorted (mtch must follow: see k_enc_sort_s)
  L.mtch = o; o += al(N * 4);               // positions by partition, radix temporary, then the matches
  L.bstart = o; o += al((size_t)nb * (HB + 1) * 4);
  L.qmid = o; o += al((size_t)L.qcap * 8);
  L.qlarge = o; o += al((size_t)L.qcap * 8);
  L.qcnt = o; o += al(16);
  L.mark = o; o += al(N / 8);
  // (the five symbol streams keep 1 KiB of slack: k_enc_rans reads them 1 KiB at a time)
  L.lit = o; o += al(N + 1024);
  L.LB = o; o += al((size_t)nb * G.TCAP * 4);
  L.TM = o; o += al((size_t)nb * G.TCAP * 4);
  L.TD = o; o += al((size_t)nb * G.TCAP * 4);
  L.tok = o; o += al((size_t)nb * G.TCAP + 1024);
  L.sym = o; o += al((size_t)nb * G.TCAP + 1024);
  L.grp = o; o += al((size_t)nb * G.TCAP + 1024);
  L.ext = o; o += al((size_t)nb * G.ECAP + 1024);
  L.offw = o; o += al((size_t)nb * (G.OCAP / 4 + 1) * 4);
  L.cnt = o; o += al((size_t)nb * 8 * 4);
  L.lanes = o; o += al((size_t)nb * lanes_per_block(G));
  L.freq = o; o += al((size_t)nb * 5 * 256 * 2);
  L.lanelen = o; o += al((size_t)nb * 5 * 64 * 4);
  L.sinfo = o; o += al((size_t)nb * 5 * 4 * 4);
  L.total = o;
  return L;
}

struct DecLayout {
  size_t lit, tok, ext, sym, grp, total;
};
DecLayout dec_layout(uint32_t nb, uint32_t blog) {
  const Geo G = geo_of(blog);
  DecLayout L;
  size_t o = 0;
  L.lit = o; o += al(((size_t)nb << blog) + 64);  // (slack: the replay reads literals 16 bytes at a time)
  L.tok = o; o += al((size_t)nb * G.TCAP);
  L.ext = o; o += al((size_t)nb * G.ECAP);
  L.sym = o; o += al((size_t)nb * G.TCAP);
  L.grp = o; o += al((size_t)nb * G.TCAP);
  L.total = o;
  return L;
}

uint32_t host_mulmod(uint32_t a, uint32_t b) {
  uint32_t r = 0;
  for (int i = 31; i >= 0; --i) {
    if ((a >> i) & 1u) r ^= b;
    b = (b >> 1) ^ ((b & 1u) ? 0xEDB88320u : 0u);
  }
  return r;
}
}  // namespace

extern "C" {

// blocks per encoder launch at most: 2^31 positions (2 GiB of input), so that no grid exceeds 2^32 threads (the
// match kernel runs one thread per position); larger calls run as several launches over the same scratch
static uint32_t enc_launch_blocks(uint32_t blog) { return 1u << (31 - blog); }
size_t zgpu_encode_scratch_bytes(uint32_t nb, uint32_t block_log) {
  return enc_layout(std::min(nb, enc_launch_blocks(block_log)), block_log, hbits_of(block_log)).total;
}
size_t zgpu_decode_scratch_bytes(uint32_t nb, uint32_t block_log) { return dec_layout(nb, block_log).total; }

int zgpu_open(int ordinal, zgpu_t** out) {
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(nullptr, ZGPU_ERR_DEVICE, "no GPU");
  if (ordinal < 0 || ordinal >= n) return ZGPU_ERR_ARG;
  zgpu_t* g = new zgpu();
  g->dev = ordinal;
  hipError_t e = hipSetDevice(ordinal);
  hipDeviceProp_t p;
  if (e == hipSuccess) e = hipGetDeviceProperties(&p, ordinal);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = noalloc_scratch_of(ordinal, &g->noalloc);
  if (e != hipSuccess) { delete g; return ZGPU_ERR_DEVICE; }
  snprintf(g->props.name, sizeof g->props.name, "%s", p.name);
  snprintf(g->props.arch, sizeof g->props.arch, "%s", p.gcnArchName);
  g->props.cu_count = p.multiProcessorCount;
  g->props.clock_khz = p.clockRate;
  g->props.hbm_bytes = p.totalGlobalMem;
  g->props.lds_per_block = (uint32_t)p.sharedMemPerBlock;
  g->props.max_group = (uint32_t)p.maxThreadsPerBlock;
  g->props.integrated = p.integrated;
  // CRC combine powers: xp[j*256 + m] = x^(8 m 256^j) mod P
  std::vector<uint32_t> xp(768);
  uint32_t step = 1u << 23;  // x^8 (bit 31 is x^0)
  for (int j = 0; j < 3; ++j) {
    uint32_t acc = 1u << 31;
    for (int m = 0; m < 256; ++m) { xp[j * 256 + m] = acc; acc = host_mulmod(acc, step); }
    step = acc;  // x^(8 * 256^(j+1))
  }
  if (hipMalloc(&g->xpow, 768 * 4) != hipSuccess ||
      hipMemcpy(g->xpow, xp.data(), 768 * 4, hipMemcpyHostToDevice) != hipSuccess) {
    delete g;
    return ZGPU_ERR_DEVICE;
  }
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_enc_hash), hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_enc_place), hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024);
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_enc_sort_l), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)sort_l_lds_bytes(20));
  (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_enc_parse), hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024);
  if (const char* pe = getenv("ZERO_PROFILE")) g->profile = atoi(pe) != 0;
  *out = g;
  return ZGPU_OK;
}

void zgpu_close(zgpu_t* g) {
  if (!g) return;
  (void)hipSetDevice(g->dev);
  (void)hipStreamSynchronize(g->stream);
  for (Lane& l : g->lanes) {
    if (l.st) (void)hipStreamSynchronize(l.st);
    for (DBuf* b : {&l.work, &l.slab, &l.pay, &l.meta})
      if (b->p) (void)hipFree(b->p);
    if (l.h_total) (void)hipHostFree(l.h_total);
    if (l.ev) (void)hipEventDestroy(l.ev);
    if (l.st) (void)hipStreamDestroy(l.st);
  }
  for (DBuf* b : {&g->enc, &g->dec, &g->aux, &g->rng, &g->hin, &g->hout})
    if (b->p) (void)hipFree(b->p);
  if (g->stage.p) (void)hipHostFree(g->stage.p);
  for (hipEvent_t e : g->pool) (void)hipEventDestroy(e);
  if (g->xpow) (void)hipFree(g->xpow);
  (void)hipStreamDestroy(g->stream);
  delete g;
}

int zgpu_props(zgpu_t* g, zgpu_props_t* out) {
  *out = g->props;
  return ZGPU_OK;
}
const char* zgpu_last_error(zgpu_t* g) { return g ? g->err.c_str() : "no context"; }
int zgpu_sync(zgpu_t* g) {
  HIPCHK(g, hipStreamSynchronize(g->stream));
  HIPCHK(g, hipDeviceSynchronize());
  return ZGPU_OK;
}

int zgpu_alloc(zgpu_t* g, size_t bytes, void** dptr) { HIPCHK(g, hipMalloc(dptr, bytes ? bytes : 1)); return ZGPU_OK; }
int zgpu_free(zgpu_t* g, void* dptr) { HIPCHK(g, hipFree(dptr)); return ZGPU_OK; }
int zgpu_alloc_host(zgpu_t* g, size_t bytes, void** hptr) {
  HIPCHK(g, hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault));
  return ZGPU_OK;
}
int zgpu_free_host(zgpu_t* g, void* hptr) { HIPCHK(g, hipHostFree(hptr)); return ZGPU_OK; }
int zgpu_upload(zgpu_t* g, void* dptr, const void* host, size_t bytes) {
  if (!bytes) return ZGPU_OK;
  HIPCHK(g, hipMemcpyAsync(dptr, host, bytes, hipMemcpyHostToDevice, g->stream));
  HIPCHK(g, hipStreamSynchronize(g->stream));
  return ZGPU_OK;
}
int zgpu_download(zgpu_t* g, void* host, const void* dptr, size_t bytes) {
  if (!bytes) return ZGPU_OK;
  HIPCHK(g, hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, g->stream));
  HIPCHK(g, hipStreamSynchronize(g->stream));
  return ZGPU_OK;
}
int zgpu_fill(zgpu_t* g, void* dptr, uint8_t value, size_t bytes, zgpu_stream_t s) {
  if (!bytes) return ZGPU_OK;
  hipStream_t st = s ? (hipStream_t)s : g->stream;
  const size_t groups = std::min<size_t>((bytes + 255) / 256, 4096);
  k_fill<<<(unsigned)groups, 256, 0, st>>>((uint8_t*)dptr, value, bytes);
  HIPCHK(g, hipGetLastError());
  return ZGPU_OK;
}

int zgpu_copy(zgpu_t* g, void* dst, const void* src, size_t bytes, zgpu_stream_t s) {
  if (!bytes) return ZGPU_OK;
  hipStream_t st = s ? (hipStream_t)s : g->stream;
  const bool aligned = (((uintptr_t)dst | (uintptr_t)src) & 15u) == 0;
  const size_t n16 = aligned ? bytes / 16 : 0;
  if (n16) {
    const size_t groups = std::min<size_t>((n16 + 255) / 256, (size_t)g->props.cu_count * 8);
    k_copy<<<(unsigned)groups, 256, 0, st>>>((uint4*)dst, (const uint4*)src, n16);
    HIPCHK(g, hipGetLastError());
  }
  if (bytes > n16 * 16)
    HIPCHK(g, hipMemcpyAsync((uint8_t*)dst + n16 * 16, (const uint8_t*)src + n16 * 16, bytes - n16 * 16,
                             hipMemcpyDeviceToDevice, st));
  return ZGPU_OK;
}

int zgpu_stream_create(zgpu_t* g, zgpu_stream_t* out) {
  hipStream_t s;
  HIPCHK(g, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = s;
  return ZGPU_OK;
}
int zgpu_stream_destroy(zgpu_t* g, zgpu_stream_t s) { HIPCHK(g, hipStreamDestroy((hipStream_t)s)); return ZGPU_OK; }
int zgpu_stream_sync(zgpu_t* g, zgpu_stream_t s) {
  HIPCHK(g, hipStreamSynchronize(s ? (hipStream_t)s : g->stream));
  return ZGPU_OK;
}
int zgpu_event_create(zgpu_t* g, void** ev) {
  hipEvent_t e;
  HIPCHK(g, hipEventCreate(&e));
  *ev = e;
  return ZGPU_OK;
}
int zgpu_event_destroy(zgpu_t* g, void* ev) { HIPCHK(g, hipEventDestroy((hipEvent_t)ev)); return ZGPU_OK; }
int zgpu_event_record(zgpu_t* g, void* ev, zgpu_stream_t s) {
  HIPCHK(g, hipEventRecord((hipEvent_t)ev, s ? (hipStream_t)s : g->stream));
  return ZGPU_OK;
}
int zgpu_event_elapsed_ms(zgpu_t* g, void* a, void* b, float* ms) {
  HIPCHK(g, hipEventSynchronize((hipEvent_t)b));
  HIPCHK(g, hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
  return ZGPU_OK;
}

int zgpu_set_profile(zgpu_t* g, int on) { g->profile = on != 0; return ZGPU_OK; }
int zgpu_stage_times(zgpu_t* g, const char** names, float* ms, int cap) {
  int k = 0;
  for (size_t i = 1; i < g->stages.size() && k < cap; ++i, ++k) {
    HIPCHK(g, hipEventSynchronize(g->stages[i].ev));
    HIPCHK(g, hipEventElapsedTime(&ms[k], g->stages[i - 1].ev, g->stages[i].ev));
    names[k] = g->stages[i].name;
  }
  return k;
}

// ---------------------------------------------------------------- encode
// the encoder's launch sequence for nb blocks of one shape, in the scratch at work (enc_layout(nb, S.blog, S.hbits))
static int encode_run(zgpu_t* g, const uint8_t* d_in, const Block* blocks, uint32_t nb, const EncShape& S, int level,
                      uint8_t* d_out, uint32_t* d_out_len, uint8_t* base, hipStream_t st, bool prof) {
  const uint32_t blog = S.blog;
  const EncLayout L = enc_layout(nb, blog, S.hbits);
  const uint32_t ncand = level <= 2 ? 2 : level <= 6 ? 4 : 8;
  const uint32_t lencap = level <= 6 ? 16 : 32;
  const uint32_t HB = 1u << S.hbits;
  uint32_t* srt = (uint32_t*)(base + L.srt);
  uint32_t* mt = (uint32_t*)(base + L.mtch);
  uint32_t* bstart = (uint32_t*)(base + L.bstart);
  SortQueues q{(uint32_t*)(base + L.qmid), (uint32_t*)(base + L.qlarge), (uint32_t*)(base + L.qcnt), L.qcap};
  TokBufs T{base + L.lit,  (uint32_t*)(base + L.LB), (uint32_t*)(base + L.TM), (uint32_t*)(base + L.TD),
            base + L.tok,  base + L.sym,              base + L.grp,              base + L.ext,
            (uint32_t*)(base + L.offw), (uint32_t*)(base + L.cnt)};
  RansBufs R{base + L.lanes, (uint16_t*)(base + L.freq), (uint32_t*)(base + L.lanelen), (uint32_t*)(base + L.sinfo)};
  const uint32_t posgroups = nb << (blog - 8);
  if (prof) stage_begin(g, st);
  k_fill<<<1, 64, 0, st>>>((uint8_t*)q.counts, 0, 16);
  const uint32_t pbits = pbits_of(blog, S.hbits);
  k_enc_hash<<<nb, 1024, hash_lds_words(S.hbits, pbits) * 4, st>>>(blocks, d_in, blog, S.hbits, pbits, mt, bstart);
  k_enc_place<<<nb << pbits, 1024, ((1u << (S.hbits - pbits)) + PLACE_CAP) * 4, st>>>(blocks, d_in, nb, blog, S.hbits,
                                                                                     S.keybits, pbits, mt, bstart, srt);
  if (prof) stage(g, st, "hash");
  k_enc_sort_s<<<(unsigned)(((size_t)nb * HB + 255) / 256), 256, 0, st>>>(blocks, nb, blog, S.hbits, srt, bstart, q);
  k_enc_sort_m<<<std::min<uint32_t>(2048, L.qcap), 256, 0, st>>>(blog, S.hbits, srt, bstart, q);
  k_enc_sort_l<<<std::min<uint32_t>(1024, L.qcap), 1024, sort_l_lds_bytes(blog), st>>>(blog, S.hbits, srt, mt, bstart,
                                                                                      q);
  if (prof) stage(g, st, "sort");
  // the match array cleared: the match kernel writes only the positions that get a match
  k_zero16<<<(unsigned)std::min<size_t>(((size_t)nb << blog) / 1024 + 1, 16384), 256, 0, st>>>((uint4*)mt,
                                                                                              ((size_t)nb << blog) / 4);
  if (prof) stage(g, st, "clear");
  k_enc_match<<<posgroups, 256, 0, st>>>(blocks, d_in, blog, ncand, lencap, srt, mt);
  if (prof) stage(g, st, "match");
  const size_t parse_lds = parse_lds_words(blog, S.walkers) * 4;
  k_enc_parse<<<nb, std::max(64u, S.walkers), parse_lds, st>>>(blocks, d_in, blog, S.walkers, lencap, mt,
                                                               (uint32_t*)(base + L.mark));
  if (prof) stage(g, st, "parse");
  k_enc_tokens<<<nb, 64, 0, st>>>(blocks, d_in, blog, lencap, mt, (uint32_t*)(base + L.mark), T);
  if (prof) stage(g, st, "tokens");
  k_enc_rans<<<nb * 5, 64, 0, st>>>(blog, T, R);
  if (prof) stage(g, st, "rans");
  k_enc_pack<<<nb, 256, 0, st>>>(blocks, blog, T, R, d_out, d_out_len);
  if (prof) stage(g, st, "pack");
  HIPCHK(g, hipGetLastError());
  return ZGPU_OK;
}

int zgpu_encode_blocks(zgpu_t* g, const uint8_t* d_in, const zgpu_block_t* d_blocks, uint32_t nb, uint32_t blog,
                       int level, uint8_t* d_out, uint32_t* d_out_len, zgpu_stream_t s) {
  if (blog < 16 || blog > 20 || level < 1 || level > 9) return fail(g, ZGPU_ERR_ARG, "block_log 16..20, level 1..9");
  if (nb == 0) return ZGPU_OK;
  hipStream_t st = s ? (hipStream_t)s : g->stream;
  const EncShape S = default_shape(blog);
  const uint32_t per = std::min(nb, enc_launch_blocks(blog));
  if (int rc = ensure(g, g->enc, enc_layout(per, blog, S.hbits).total)) return rc;
  for (uint32_t b0 = 0; b0 < nb; b0 += per)  // (in order on one stream: each launch's scratch is free for the next)
    if (int rc = encode_run(g, d_in, (const Block*)d_blocks + b0, std::min(per, nb - b0), S, level,
                            d_out + ((size_t)b0 << blog), d_out_len + b0, (uint8_t*)g->enc.p, st, b0 == 0))
      return rc;
  return ZGPU_OK;
}

// one stream section (u8 kind, u32 len, body) of n symbols by the block encoder's own kernels: the symbols are the
// literal stream of a single block whose other streams are empty, k_enc_rans normalises and encodes them, k_enc_pack
// lays the payload out, and the section is the payload's first one. The block is 2^blog with 2^blog >= 2n + 4096,
// so the payload always fits and is never "stored".
int zgpu_encode_stream(zgpu_t* g, const uint8_t* sym, uint32_t n, uint8_t* out, uint32_t* out_len) {
  if (n > (1u << 19)) return fail(g, ZGPU_ERR_ARG, "a stream holds at most 2^19 symbols");
  uint32_t blog = 16;
  while (blog < 20 && (1ull << blog) < 2ull * n + 4096) ++blog;
  const Geo G = geo_of(blog);
  const EncLayout L = enc_layout(1, blog, hbits_of(blog));
  if (int rc = ensure(g, g->enc, L.total)) return rc;
  if (int rc = ensure(g, g->aux, al(G.B) + al(64) + al(32))) return rc;
  g->last_len[0] = g->last_len[1] = 0;
  uint8_t* base = (uint8_t*)g->enc.p;
  uint8_t* d_out = (uint8_t*)g->aux.p;
  uint32_t* d_len = (uint32_t*)(d_out + al(G.B));
  zgpu_block_t* d_tab = (zgpu_block_t*)((uint8_t*)d_len + al(64));
  zgpu_block_t tab{};
  tab.raw_len = G.B;
  const uint32_t cnt[8] = {0, n, 0, 0, 0, 0, 0, 0};  // tokens, literals, ext bytes, offset-bit bytes, groups
  hipStream_t st = g->stream;
  TokBufs T{base + L.lit,  (uint32_t*)(base + L.LB), (uint32_t*)(base + L.TM), (uint32_t*)(base + L.TD),
            base + L.tok,  base + L.sym,              base + L.grp,              base + L.ext,
            (uint32_t*)(base + L.offw), (uint32_t*)(base + L.cnt)};
  RansBufs R{base + L.lanes, (uint16_t*)(base + L.freq), (uint32_t*)(base + L.lanelen), (uint32_t*)(base + L.sinfo)};
  if (n) HIPCHK(g, hipMemcpyAsync(T.lit, sym, n, hipMemcpyHostToDevice, st));
  HIPCHK(g, hipMemcpyAsync(T.cnt, cnt, sizeof cnt, hipMemcpyHostToDevice, st));
  HIPCHK(g, hipMemcpyAsync(d_tab, &tab, sizeof tab, hipMemcpyHostToDevice, st));
  k_enc_rans<<<5, 64, 0, st>>>(blog, T, R);
  k_enc_pack<<<1, 256, 0, st>>>((const Block*)d_tab, blog, T, R, d_out, d_len);
  HIPCHK(g, hipGetLastError());
  uint32_t total = 0;
  HIPCHK(g, hipMemcpyAsync(&total, d_len, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(g, hipStreamSynchronize(st));
  if (total == ZGPU_STORED || total < 20 + 5 * 5) return fail(g, ZGPU_ERR_DEVICE, "stream encoder: impossible payload");
  uint8_t hdr[5];
  HIPCHK(g, hipMemcpy(hdr, d_out + 20, 5, hipMemcpyDeviceToHost));
  const uint32_t len = (uint32_t)hdr[1] | (uint32_t)hdr[2] << 8 | (uint32_t)hdr[3] << 16 | (uint32_t)hdr[4] << 24;
  if (20ull + 5 + len > total) return fail(g, ZGPU_ERR_DEVICE, "stream encoder: section past the payload");
  HIPCHK(g, hipMemcpy(out, d_out + 20, 5 + (size_t)len, hipMemcpyDeviceToHost));
  *out_len = 5 + len;
  return ZGPU_OK;
}

// ---------------------------------------------------------------- decode
static int decode_run(zgpu_t* g, const uint8_t* d_comp, const Block* blocks, uint32_t nb, uint32_t blog,
                      uint8_t* d_out, uint32_t* d_err, uint8_t* base, hipStream_t st, bool prof) {
  const DecLayout L = dec_layout(nb, blog);
  if (prof) stage_begin(g, st);
  k_fill<<<(unsigned)std::min<size_t>((nb * 4 + 255) / 256, 1024), 256, 0, st>>>((uint8_t*)d_err, 0, (size_t)nb * 4);
  k_dec_streams<<<nb * 5, 64, 0, st>>>(blocks, d_comp, blog, base + L.lit, base + L.tok, base + L.ext, base + L.sym,
                                       base + L.grp, d_err);
  if (prof) stage(g, st, "streams");
  k_dec_lz<<<nb, 64, 0, st>>>(blocks, d_comp, blog, base + L.lit, base + L.tok, base + L.ext, base + L.sym,
                              base + L.grp, d_out, d_err);
  if (prof) stage(g, st, "lz");
  HIPCHK(g, hipGetLastError());
  return ZGPU_OK;
}

int zgpu_decode_blocks(zgpu_t* g, const uint8_t* d_comp, const zgpu_block_t* d_blocks, uint32_t nb, uint32_t blog,
                       uint8_t* d_out, uint32_t* d_err, zgpu_stream_t s) {
  if (blog < 16 || blog > 20) return fail(g, ZGPU_ERR_ARG, "block_log 16..20");
  if (nb == 0) return ZGPU_OK;
  hipStream_t st = s ? (hipStream_t)s : g->stream;
  if (int rc = ensure(g, g->dec, dec_layout(nb, blog).total)) return rc;
  return decode_run(g, d_comp, (const Block*)d_blocks, nb, blog, d_out, d_err, (uint8_t*)g->dec.p, st, true);
}

// zgpu_decode_blocks with the decoder's scratch in the fixed array `scratch` (through its device address): no
// allocation, no event, no library call that allocates. Calls that overlap on one GPU share that array and race.
int zgpu_decode_blocks_noalloc(zgpu_t* g, const uint8_t* d_comp, const zgpu_block_t* d_blocks, uint32_t nb,
                               uint32_t blog, uint8_t* d_out, uint32_t* d_err, zgpu_stream_t s) {
  if (blog < 16 || blog > 20) return fail(g, ZGPU_ERR_ARG, "block_log 16..20");
  if (nb == 0) return ZGPU_OK;
  const size_t need = dec_layout(nb, blog).total;
  if (need > (size_t)ZGPU_NOALLOC_SCRATCH_BYTES) {
    char m[160];
    snprintf(m, sizeof m, "decode_blocks_noalloc: %u blocks of 2^%u need %zu bytes of scratch, more than the %llu it has",
             nb, blog, need, (unsigned long long)ZGPU_NOALLOC_SCRATCH_BYTES);
    return fail(g, ZGPU_ERR_ARG, m);
  }
  if (!g->noalloc) return fail(g, ZGPU_ERR_DEVICE, "decode_blocks_noalloc: no scratch address on this device");
  hipStream_t st = s ? (hipStream_t)s : g->stream;
  return decode_run(g, d_comp, (const Block*)d_blocks, nb, blog, d_out, d_err, (uint8_t*)g->noalloc, st, false);
}

int zgpu_crc32_blocks(zgpu_t* g, const uint8_t* d_data, const zgpu_block_t* d_blocks, uint32_t nb, int use_dst,
                      uint32_t* d_crc, zgpu_stream_t s) {
  if (nb == 0) return ZGPU_OK;
  hipStream_t st = s ? (hipStream_t)s : g->stream;
  k_crc32<<<(nb + 3) / 4, 256, 0, st>>>((const Block*)d_blocks, nb, d_data, use_dst, g->xpow, d_crc);
  HIPCHK(g, hipGetLastError());
  return ZGPU_OK;
}

// ---------------------------------------------------------------- host-buffer forms
// A batch moves in chunks of blocks over the two lanes (streams), chunk c on lane c % 2: while one lane's kernels
// run, the other lane's copies move over PCIe, so upload, kernels and download overlap instead of taking turns. The
// whole batch's input (compress) or output (decompress) stays on the device, for zgpu_crc32_ranges.
namespace {
// blocks per chunk of a batch of nb blocks of 2^blog: few chunks, so that each launch still fills the GPU (a
// per-block kernel runs one wavefront or workgroup per block), but at least two from 512 blocks on, so that copies
// have kernels to hide behind; at most 1 GiB of input per chunk, which bounds the lanes' scratch
uint32_t chunk_blocks(uint32_t nb, uint32_t blog) {
  const uint32_t chunks = std::max<uint32_t>(1, std::min<uint32_t>(4, nb / 256));
  return std::min<uint32_t>((nb + chunks - 1) / chunks, (1u << 30) >> blog);
}
}  // namespace

// the payload offsets of a chunk's blocks: goff[b] = the sizes of its compressed blocks before b (stored and bad
// blocks take no room), *total = all of them; one workgroup, each thread a run of ceil(nb / 1024) blocks
__global__ __launch_bounds__(1024) void k_payload_offsets(const uint32_t* __restrict__ len, uint32_t nb,
                                                          uint64_t* __restrict__ goff, uint64_t* __restrict__ total) {
  __shared__ uint64_t part[1024];
  const uint32_t t = threadIdx.x, per = (nb + 1023) / 1024, b0 = t * per;
  uint64_t s = 0;
  for (uint32_t b = b0; b < min(nb, b0 + per); ++b) s += len[b] >= BAD_BLOCK ? 0u : len[b];
  part[t] = s;
  __syncthreads();
  for (uint32_t d = 1; d < 1024; d <<= 1) {  // inclusive scan of the 1024 partial sums
    const uint64_t v = t >= d ? part[t - d] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  uint64_t o = part[t] - s;
  for (uint32_t b = b0; b < min(nb, b0 + per); ++b) {
    goff[b] = o;
    o += len[b] >= BAD_BLOCK ? 0u : len[b];
  }
  if (t == 1023) *total = part[1023];
}

// block b's payload (len[b] bytes at slab + (b << blog); nothing when stored) to pay + goff[b]
__global__ __launch_bounds__(256) void k_pack_payloads(const uint8_t* __restrict__ slab, uint32_t blog,
                                                       const uint32_t* __restrict__ len, const uint64_t* __restrict__ goff,
                                                       uint8_t* __restrict__ pay) {
  const uint32_t b = blockIdx.x, n = len[b];
  if (n >= BAD_BLOCK) return;
  const uint8_t* s = slab + ((size_t)b << blog);
  uint8_t* d = pay + goff[b];
  const uint32_t full = n & ~15u;
  for (uint32_t i = threadIdx.x * 16; i < full; i += 256 * 16) {
    uint4 v;
    __builtin_memcpy(&v, s + i, 16);
    __builtin_memcpy(d + i, &v, 16);
  }
  if (full + threadIdx.x < n) d[full + threadIdx.x] = s[full + threadIdx.x];
}

static int lanes_ready(zgpu_t* g) {
  for (Lane& l : g->lanes) {
    if (!l.st) HIPCHK(g, hipStreamCreateWithFlags(&l.st, hipStreamNonBlocking));
    if (!l.ev) HIPCHK(g, hipEventCreateWithFlags(&l.ev, hipEventDisableTiming));
    if (!l.h_total) HIPCHK(g, hipHostMalloc((void**)&l.h_total, 64, hipHostMallocDefault));
  }
  return ZGPU_OK;
}

int zgpu_compress(zgpu_t* g, const uint8_t* in, uint64_t n, uint32_t blog, int level, uint8_t* out, uint64_t out_cap,
                  uint32_t* out_len, uint32_t* crc) {
  if (blog < 16 || blog > 20 || level < 1 || level > 9) return fail(g, ZGPU_ERR_ARG, "block_log 16..20, level 1..9");
  const uint64_t B = 1ull << blog;
  const uint32_t nb = (uint32_t)((n + B - 1) / B);
  g->last_len[0] = g->last_len[1] = 0;
  if (nb == 0) return ZGPU_OK;
  if (int rc = lanes_ready(g)) return rc;
  const uint32_t per = chunk_blocks(nb, blog), nchunks = (nb + per - 1) / per;
  const EncShape S = default_shape(blog);
  // device: the whole input; per lane: scratch, slabs, packed payloads, the chunk's table, lengths, CRCs, offsets
  if (int rc = ensure(g, g->hin, al(n) + 64)) return rc;
  const size_t meta = al((size_t)per * 32) + 2 * al((size_t)per * 4) + al((size_t)per * 8) + al(8);
  for (Lane& l : g->lanes) {
    if (int rc = ensure(g, l.work, enc_layout(per, blog, S.hbits).total)) return rc;
    if (int rc = ensure(g, l.slab, (size_t)per << blog)) return rc;
    if (int rc = ensure(g, l.pay, (size_t)per << blog)) return rc;
    if (int rc = ensure(g, l.meta, meta)) return rc;
  }
  uint8_t* d_in = (uint8_t*)g->hin.p;
  // pinned: the block table, then the lengths and CRCs as they come back (copied to the caller's arrays at the end)
  if (int rc = ensure_pinned(g, g->stage, al((size_t)nb * 32) + 2 * al((size_t)nb * 4))) return rc;
  zgpu_block_t* tab = (zgpu_block_t*)g->stage.p;
  uint32_t* h_len = (uint32_t*)((uint8_t*)g->stage.p + al((size_t)nb * 32));
  uint32_t* h_crc = (uint32_t*)((uint8_t*)h_len + al((size_t)nb * 4));
  for (uint32_t b = 0; b < nb; ++b) {
    tab[b] = zgpu_block_t{};
    tab[b].src = (uint64_t)(b % per) * B;  // (within its chunk's part of the input)
    tab[b].raw_len = (uint32_t)std::min<uint64_t>(B, n - (uint64_t)b * B);
  }
  auto front = [&](uint32_t c) -> int {  // chunk c: upload, kernels, its lengths and CRCs back
    Lane& l = g->lanes[c & 1];
    const uint32_t b0 = c * per, k = std::min(per, nb - b0);
    const uint64_t o = (uint64_t)b0 * B, bytes = std::min<uint64_t>((uint64_t)k * B, n - o);
    uint8_t* m = (uint8_t*)l.meta.p;
    zgpu_block_t* d_tab = (zgpu_block_t*)m;
    uint32_t* d_len = (uint32_t*)(m + al((size_t)per * 32));
    uint32_t* d_crc = (uint32_t*)((uint8_t*)d_len + al((size_t)per * 4));
    uint64_t* d_goff = (uint64_t*)((uint8_t*)d_crc + al((size_t)per * 4));
    uint64_t* d_tot = (uint64_t*)((uint8_t*)d_goff + al((size_t)per * 8));
    HIPCHK(g, hipMemcpyAsync(d_in + o, in + o, bytes, hipMemcpyHostToDevice, l.st));
    HIPCHK(g, hipMemcpyAsync(d_tab, tab + b0, (size_t)k * 32, hipMemcpyHostToDevice, l.st));
    if (int rc = encode_run(g, d_in + o, (const Block*)d_tab, k, S, level, (uint8_t*)l.slab.p, d_len,
                            (uint8_t*)l.work.p, l.st, c == 0))
      return rc;
    k_payload_offsets<<<1, 1024, 0, l.st>>>(d_len, k, d_goff, d_tot);
    k_pack_payloads<<<k, 256, 0, l.st>>>((const uint8_t*)l.slab.p, blog, d_len, d_goff, (uint8_t*)l.pay.p);
    HIPCHK(g, hipGetLastError());
    if (crc) {
      k_crc32<<<(k + 3) / 4, 256, 0, l.st>>>((const Block*)d_tab, k, d_in + o, 0, g->xpow, d_crc);
      HIPCHK(g, hipGetLastError());
      HIPCHK(g, hipMemcpyAsync(h_crc + b0, d_crc, (size_t)k * 4, hipMemcpyDeviceToHost, l.st));
    }
    HIPCHK(g, hipMemcpyAsync(h_len + b0, d_len, (size_t)k * 4, hipMemcpyDeviceToHost, l.st));
    HIPCHK(g, hipMemcpyAsync(l.h_total, d_tot, 8, hipMemcpyDeviceToHost, l.st));
    HIPCHK(g, hipEventRecord(l.ev, l.st));
    return ZGPU_OK;
  };
  if (int rc = front(0)) return rc;
  if (nchunks > 1)
    if (int rc = front(1)) return rc;
  uint64_t base = 0;
  for (uint32_t c = 0; c < nchunks; ++c) {
    Lane& l = g->lanes[c & 1];
    HIPCHK(g, hipEventSynchronize(l.ev));  // (the other lane's chunk keeps the GPU busy meanwhile)
    const uint64_t total = *l.h_total;
    if (total > out_cap - base) return fail(g, ZGPU_ERR_ARG, "compress: the output buffer is too small");
    if (total) HIPCHK(g, hipMemcpyAsync(out + base, l.pay.p, total, hipMemcpyDeviceToHost, l.st));
    base += total;
    if (c + 2 < nchunks)
      if (int rc = front(c + 2)) return rc;  // (after this lane's download: the stream keeps them in order)
  }
  for (Lane& l : g->lanes) HIPCHK(g, hipStreamSynchronize(l.st));
  memcpy(out_len, h_len, (size_t)nb * 4);
  if (crc) memcpy(crc, h_crc, (size_t)nb * 4);
  g->last_data[0] = d_in;
  g->last_len[0] = n;
  for (uint32_t b = 0; b < nb; ++b)
    if (out_len[b] != ZGPU_STORED && out_len[b] >= tab[b].raw_len)
      return fail(g, ZGPU_ERR_DEVICE, "GPU encoder produced an impossible block size");
  return ZGPU_OK;
}

int zgpu_decompress(zgpu_t* g, const uint8_t* comp, const uint32_t* comp_len, const uint8_t* method,
                    const uint32_t* raw_len, uint32_t nb, uint32_t blog, uint8_t* out, uint32_t* err, uint32_t* crc) {
  if (blog < 16 || blog > 20) return fail(g, ZGPU_ERR_ARG, "block_log 16..20");
  g->last_len[0] = g->last_len[1] = 0;
  if (nb == 0) return ZGPU_OK;
  const uint64_t B = 1ull << blog;
  std::vector<uint64_t> co(nb + 1), ro(nb + 1);
  co[0] = ro[0] = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    if (raw_len[b] > B) return fail(g, ZGPU_ERR_DAMAGED, "block larger than the archive's block size");
    co[b + 1] = co[b] + comp_len[b];
    ro[b + 1] = ro[b] + raw_len[b];
  }
  if (int rc = lanes_ready(g)) return rc;
  const uint32_t per = chunk_blocks(nb, blog), nchunks = (nb + per - 1) / per;
  if (int rc = ensure(g, g->hin, al(co[nb]) + 64)) return rc;
  if (int rc = ensure(g, g->hout, al(ro[nb]) + 64)) return rc;
  for (Lane& l : g->lanes) {
    if (int rc = ensure(g, l.work, dec_layout(per, blog).total)) return rc;
    if (int rc = ensure(g, l.meta, al((size_t)per * 32) + 2 * al((size_t)per * 4))) return rc;
  }
  uint8_t* d_comp = (uint8_t*)g->hin.p;
  uint8_t* d_out = (uint8_t*)g->hout.p;
  // pinned: the block table, then the error bits and CRCs as they come back (copied out at the end)
  if (int rc = ensure_pinned(g, g->stage, al((size_t)nb * 32) + 2 * al((size_t)nb * 4))) return rc;
  zgpu_block_t* tab = (zgpu_block_t*)g->stage.p;
  uint32_t* h_err = (uint32_t*)((uint8_t*)g->stage.p + al((size_t)nb * 32));
  uint32_t* h_crc = (uint32_t*)((uint8_t*)h_err + al((size_t)nb * 4));
  for (uint32_t c = 0; c < nchunks; ++c) {  // every size is known: the chunks are queued at once
    Lane& l = g->lanes[c & 1];
    const uint32_t b0 = c * per, k = std::min(per, nb - b0);
    for (uint32_t b = b0; b < b0 + k; ++b) {
      tab[b] = zgpu_block_t{};
      tab[b].src = co[b] - co[b0];
      tab[b].dst = ro[b] - ro[b0];
      tab[b].raw_len = raw_len[b];
      tab[b].comp_len = comp_len[b];
      tab[b].method = method[b];
      tab[b].level = 0;
    }
    uint8_t* m = (uint8_t*)l.meta.p;
    zgpu_block_t* d_tab = (zgpu_block_t*)m;
    uint32_t* d_err = (uint32_t*)(m + al((size_t)per * 32));
    uint32_t* d_crc = (uint32_t*)((uint8_t*)d_err + al((size_t)per * 4));
    HIPCHK(g, hipMemcpyAsync(d_comp + co[b0], comp + co[b0], co[b0 + k] - co[b0], hipMemcpyHostToDevice, l.st));
    HIPCHK(g, hipMemcpyAsync(d_tab, tab + b0, (size_t)k * 32, hipMemcpyHostToDevice, l.st));
    if (int rc = decode_run(g, d_comp + co[b0], (const Block*)d_tab, k, blog, d_out + ro[b0], d_err,
                            (uint8_t*)l.work.p, l.st, c == 0))
      return rc;
    if (crc) {
      k_crc32<<<(k + 3) / 4, 256, 0, l.st>>>((const Block*)d_tab, k, d_out + ro[b0], 1, g->xpow, d_crc);
      HIPCHK(g, hipGetLastError());
      HIPCHK(g, hipMemcpyAsync(h_crc + b0, d_crc, (size_t)k * 4, hipMemcpyDeviceToHost, l.st));
    }
    HIPCHK(g, hipMemcpyAsync(h_err + b0, d_err, (size_t)k * 4, hipMemcpyDeviceToHost, l.st));
    HIPCHK(g, hipMemcpyAsync(out + ro[b0], d_out + ro[b0], ro[b0 + k] - ro[b0], hipMemcpyDeviceToHost, l.st));
  }
  for (Lane& l : g->lanes) HIPCHK(g, hipStreamSynchronize(l.st));
  memcpy(err, h_err, (size_t)nb * 4);
  if (crc) memcpy(crc, h_crc, (size_t)nb * 4);
  g->last_data[1] = d_out;
  g->last_len[1] = ro[nb];
  for (uint32_t b = 0; b < nb; ++b)
    if (err[b]) {
      char msg[96];
      snprintf(msg, sizeof msg, "block %u damaged (GPU decoder, code 0x%x)", b, err[b]);
      return fail(g, ZGPU_ERR_DAMAGED, msg);
    }
  return ZGPU_OK;
}

int zgpu_crc32(zgpu_t* g, const uint8_t* data, uint64_t n, uint32_t blog, uint32_t* crc) {
  if (blog < 16 || blog > 20) return fail(g, ZGPU_ERR_ARG, "block_log 16..20");
  const uint64_t B = 1ull << blog;
  const uint32_t nb = (uint32_t)((n + B - 1) / B);
  if (nb == 0) return ZGPU_OK;
  std::vector<zgpu_block_t> tab(nb);
  for (uint32_t b = 0; b < nb; ++b) {
    tab[b] = zgpu_block_t{};
    tab[b].src = (uint64_t)b * B;
    tab[b].raw_len = (uint32_t)std::min<uint64_t>(B, n - (uint64_t)b * B);
  }
  g->last_len[0] = g->last_len[1] = 0;
  if (int rc = ensure(g, g->aux, al(n) + al((size_t)nb * 4) + al((size_t)nb * 32))) return rc;
  uint8_t* d_in = (uint8_t*)g->aux.p;
  uint32_t* d_crc = (uint32_t*)(d_in + al(n));
  zgpu_block_t* d_tab = (zgpu_block_t*)((uint8_t*)d_crc + al((size_t)nb * 4));
  HIPCHK(g, hipMemcpyAsync(d_in, data, n, hipMemcpyHostToDevice, g->stream));
  HIPCHK(g, hipMemcpyAsync(d_tab, tab.data(), (size_t)nb * 32, hipMemcpyHostToDevice, g->stream));
  if (int rc = zgpu_crc32_blocks(g, d_in, d_tab, nb, 0, d_crc, nullptr)) return rc;
  HIPCHK(g, hipMemcpyAsync(crc, d_crc, (size_t)nb * 4, hipMemcpyDeviceToHost, g->stream));
  HIPCHK(g, hipStreamSynchronize(g->stream));
  return ZGPU_OK;
}

// ---------------------------------------------------------------- CRCs of byte ranges (files in the solid stream)
namespace {
uint32_t crc_mulmod(uint32_t a, uint32_t b) {  // reflected: bit 31 is x^0
  uint32_t r = 0;
  for (int i = 31; i >= 0; --i) {
    if ((a >> i) & 1u) r ^= b;
    b = (b >> 1) ^ ((b & 1u) ? 0xEDB88320u : 0u);
  }
  return r;
}
uint32_t crc_combine(uint32_t a, uint32_t b, uint64_t len_b) {  // crc(A||B) from crc(A), crc(B), |B|
  uint32_t acc = 1u << 31, sq = 1u << 23;                      // x^0, x^8
  for (uint64_t e = len_b; e; e >>= 1) {
    if (e & 1u) acc = crc_mulmod(acc, sq);
    sq = crc_mulmod(sq, sq);
  }
  return crc_mulmod(a, acc) ^ b;
}
constexpr uint64_t RANGE_PIECE = 1u << 18;  // ranges go to the GPU in pieces of at most this (one wavefront each)
}  // namespace

int zgpu_crc32_ranges(zgpu_t* g, int which, uint32_t n, const uint64_t* off, const uint64_t* len, uint32_t* crc) {
  if (which != 0 && which != 1) return fail(g, ZGPU_ERR_ARG, "which: 0 (last compress input) or 1 (last decompress output)");
  if (n == 0) return ZGPU_OK;
  const uint64_t total = g->last_len[which];
  std::vector<zgpu_block_t> tab;
  for (uint32_t i = 0; i < n; ++i) {
    if (off[i] > total || len[i] > total - off[i])
      return fail(g, ZGPU_ERR_ARG, "range outside the last call's data (or no such call since the last other one)");
    for (uint64_t o = 0; o < len[i]; o += RANGE_PIECE) {
      zgpu_block_t t{};
      t.src = off[i] + o;
      t.raw_len = (uint32_t)std::min<uint64_t>(RANGE_PIECE, len[i] - o);
      tab.push_back(t);
    }
  }
  const uint32_t np = (uint32_t)tab.size();
  std::vector<uint32_t> pc(np);
  if (np) {
    if (int rc = ensure(g, g->rng, al((size_t)np * 32) + al((size_t)np * 4))) return rc;
    zgpu_block_t* d_tab = (zgpu_block_t*)g->rng.p;
    uint32_t* d_crc = (uint32_t*)((uint8_t*)g->rng.p + al((size_t)np * 32));
    HIPCHK(g, hipMemcpyAsync(d_tab, tab.data(), (size_t)np * 32, hipMemcpyHostToDevice, g->stream));
    if (int rc = zgpu_crc32_blocks(g, g->last_data[which], d_tab, np, 0, d_crc, nullptr)) return rc;
    HIPCHK(g, hipMemcpyAsync(pc.data(), d_crc, (size_t)np * 4, hipMemcpyDeviceToHost, g->stream));
    HIPCHK(g, hipStreamSynchronize(g->stream));
  }
  uint32_t k = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t c = 0;
    for (uint64_t o = 0; o < len[i]; o += RANGE_PIECE, ++k) c = crc_combine(c, pc[k], tab[k].raw_len);
    crc[i] = c;
  }
  return ZGPU_OK;
}

}  // extern "C"

#include "plan.hip"
#include "serve.hip"
