// C ABI of the MI355X summed-area-table library: argument validation,
// workspaces, two-pass compositions and the batch planner. Kernels live in
// k_*.hip; nothing here falls back to host arithmetic.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>
#include <vector>

#include "sat_host.hpp"

namespace sathost {

sat_status launch_fam_sat(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_sat_rev(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_rows(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_rows_t(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_cols(const Launch& L, void* ws, size_t b, hipStream_t st);
sat_status launch_fam_serial_t(const Launch& L, void* ws, size_t b, hipStream_t st);

size_t dtype_size(int dt) {
  switch (dt) {
    case SAT_U8:
    case SAT_I8: return 1;
    case SAT_U16:
    case SAT_I16: return 2;
    case SAT_U32:
    case SAT_I32:
    case SAT_F32: return 4;
    case SAT_F64: return 8;
    default: return 0;
  }
}

static uint64_t tiles_of(const Geometry& g, uint32_t width, uint32_t height, uint32_t planes) {
  if (g.th == 0 || g.tw == 0) return 0;
  return (uint64_t)((width + g.tw - 1) / g.tw) * ((height + g.th - 1) / g.th) * planes;
}

size_t carry_bytes(const Launch& L) {
  const Geometry g = geometry(L.family, dtype_size(L.tout), L.cfg, dtype_size(L.tin));
  const uint64_t n = tiles_of(g, L.width, L.height, L.planes);
  uint64_t words = n * (uint64_t)(g.vwords + g.hwords);
  if (g.gwords && n) {
    // the second vertical carry level of tall images (as launch_op lays it out)
    const uint64_t nTx = (L.width + g.tw - 1) / g.tw, nTy = (L.height + g.th - 1) / g.th;
    const uint32_t G = vgroup_len(nTx, nTy);
    if (G) words += (uint64_t)L.planes * ((nTy + G - 1) / G) * nTx * (uint64_t)g.gwords;
  }
  return (size_t)(satdev::WS_HEADER_BYTES + words * 8);
}

static bool tiled_by_cfg(int family) { return family == FAM_SAT || family == FAM_SAT_REV; }

// Workspace bytes that fit the launch in every tile configuration, so that a
// caller-sized workspace stays valid whatever the tuning table picks.
static size_t carry_bytes_any_cfg(Launch L) {
  if (!tiled_by_cfg(L.family)) return carry_bytes(L);
  size_t m = 0;
  for (int c = 0; c < TILE_CFGS; ++c) {
    L.cfg = c;
    m = std::max(m, carry_bytes(L));
  }
  return m;
}

// ---------------------------------------------------------------------------
// Tile-configuration table: the autotuner's choice per input width (1, 4 or
// 8 bytes), output width (4 or 8 bytes) and shape class; 0 means unset.
// ---------------------------------------------------------------------------
constexpr int TUNE_CX = 9, TUNE_CY = 12, TUNE_CP = 4;
static std::atomic<int> g_tune[3][2][TUNE_CX][TUNE_CY][TUNE_CP];  // cfg + 1
static std::atomic<int> g_tune_batch[2];                          // cfg + 1

static int in_index(int in_bytes) { return in_bytes == 1 ? 0 : in_bytes == 4 ? 1 : in_bytes == 8 ? 2 : -1; }
static bool widths_ok(int in_bytes, int out_bytes) {
  return in_index(in_bytes) >= 0 && (out_bytes == 4 || out_bytes == 8) && in_bytes <= out_bytes;
}

// The autotuner's shipped table (sat_amd/tuned_gfx950.json, turned into
// tuned_table.inc by the build) is the table's content when the library is
// loaded; sat_tune_set overrides an entry, -1 returns it to unset.
struct ShippedTune {
  int in_bytes, out_bytes, cx, cy, cp, cfg;
};
struct ShippedBatch {
  int out_bytes, cfg;
};
#include "tuned_table.inc"

static bool load_shipped_table() {
  for (const ShippedTune& e : k_shipped_tune) {
    if (e.cfg < 0 || e.cfg >= TILE_CFGS || !widths_ok(e.in_bytes, e.out_bytes)) continue;
    if (e.cx < 0 || e.cx >= TUNE_CX || e.cy < 0 || e.cy >= TUNE_CY || e.cp < 0 || e.cp >= TUNE_CP) continue;
    g_tune[in_index(e.in_bytes)][e.out_bytes == 8 ? 1 : 0][e.cx][e.cy][e.cp].store(e.cfg + 1,
                                                                                   std::memory_order_relaxed);
  }
  for (const ShippedBatch& e : k_shipped_batch)
    if (e.cfg >= 0 && e.cfg < BATCH_CFGS && (e.out_bytes == 4 || e.out_bytes == 8))
      g_tune_batch[e.out_bytes == 8 ? 1 : 0].store(e.cfg + 1, std::memory_order_relaxed);
  return true;
}
[[maybe_unused]] static const bool g_shipped_loaded = load_shipped_table();

static int floor_log2(uint64_t v) {
  int r = 0;
  while (v > 1) {
    v >>= 1;
    ++r;
  }
  return r;
}

// class_x: tile columns (log2), class_y: tile rows of one plane (log2),
// class_p: planes 1 / 2-3 / 4-15 / 16+
static void tune_class(uint32_t width, uint32_t height, uint32_t planes, int* cx, int* cy, int* cp) {
  const uint64_t tx = ((uint64_t)width + 255) / 256, ty = ((uint64_t)height + 63) / 64;
  *cx = std::min(floor_log2(tx ? tx : 1), TUNE_CX - 1);
  *cy = std::min(floor_log2(ty ? ty : 1), TUNE_CY - 1);
  *cp = planes <= 1 ? 0 : planes < 4 ? 1 : planes < 16 ? 2 : 3;
}

// Built-in choice: narrow images (fewer than 8 tile columns, so no XCD column
// groups) with at least 4096 rows take 8-wave tiles of twice the height,
// which halve the vertical carry words per pixel; everything else the 4-wave
// tile. The autotuner's table overrides it per class.
static int builtin_cfg(int cx, int cy, int cp) {
  (void)cp;
  return (cx < 3 && cy >= 6) ? 1 : 0;
}

static bool class_ok(int cx, int cy, int cp) {
  return cx >= 0 && cx < TUNE_CX && cy >= 0 && cy < TUNE_CY && cp >= 0 && cp < TUNE_CP;
}

// An unset (input width, output width) class takes the entry of the
// same-width pair (input width = output width), then the built-in rule.
static int tune_lookup(int in_bytes, int out_bytes, int cx, int cy, int cp) {
  const int o = out_bytes == 8 ? 1 : 0;
  int v = g_tune[in_index(in_bytes)][o][cx][cy][cp].load(std::memory_order_relaxed);
  if (v <= 0 && in_bytes != out_bytes) v = g_tune[in_index(out_bytes)][o][cx][cy][cp].load(std::memory_order_relaxed);
  return v > 0 ? v - 1 : builtin_cfg(cx, cy, cp);
}

static int tune_pick(size_t in_size, size_t out_size, uint32_t width, uint32_t height, uint32_t planes) {
  int cx, cy, cp;
  tune_class(width, height, planes, &cx, &cy, &cp);
  return tune_lookup((int)in_size, out_size == 8 ? 8 : 4, cx, cy, cp);
}

static int batch_cfg_pick(size_t out_size) {
  const int v = g_tune_batch[out_size == 8 ? 1 : 0].load(std::memory_order_relaxed);
  return v > 0 ? v - 1 : 0;
}

sat_status launch_family(const Launch& L, void* ws, size_t ws_bytes, hipStream_t st) {
  switch (L.family) {
    case FAM_SAT: return launch_fam_sat(L, ws, ws_bytes, st);
    case FAM_SAT_REV: return launch_fam_sat_rev(L, ws, ws_bytes, st);
    case FAM_ROWS: return launch_fam_rows(L, ws, ws_bytes, st);
    case FAM_ROWS_T: return launch_fam_rows_t(L, ws, ws_bytes, st);
    case FAM_COLS: return launch_fam_cols(L, ws, ws_bytes, st);
    case FAM_SERIAL_T: return launch_fam_serial_t(L, ws, ws_bytes, st);
    default: return SAT_ERR_INVALID;
  }
}

int device_cu_count() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) return 1;
  return n;
}

int device_xcd_count() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1;
  if (dev >= 0 && dev < 64) {
    const int c = cache[dev].load(std::memory_order_relaxed);
    if (c > 0) return c;
  }
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || n < 1) n = 1;
  if (dev >= 0 && dev < 64) cache[dev].store(n, std::memory_order_relaxed);
  return n;
}

// ---------------------------------------------------------------------------
// Library-managed workspaces, one set per (device, stream), grown on demand.
// Carry words are validated by their launch tag, which is only sound if the
// carry buffer never holds anything but zeros or tagged words: the buffer
// holding the launch header and carries is therefore kept apart from the
// buffer holding two-pass temporaries (raw pixel sums could otherwise pass
// for a tag). A caller-provided workspace holds both, so its carry area is
// cleared by a fill kernel before every call (its header's fault word is
// kept until sat_workspace_fault reads it).
//
// A managed buffer is never freed while the library runs (only by
// sat_release_workspaces): when a call needs a larger one, the old buffer is
// retired and a new one, at least twice as large, takes its place. Launches
// already enqueued, launches another host thread is about to enqueue with the
// old pointer, and graphs captured with it therefore keep valid memory.
// Epochs are counted per buffer, so a retired buffer's tags never repeat.
// ---------------------------------------------------------------------------
struct Buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  uint32_t epoch = 0;  // host epoch counter (carry buffers)
};
struct ManagedWs {
  Buffer* cur = nullptr;
  std::vector<Buffer*> retired;
};
static std::mutex g_ws_mu;
static std::map<std::tuple<int, uintptr_t, int>, ManagedWs> g_ws;

enum WsKind : int { WS_CARRY = 0, WS_TMP = 1 };

static sat_status managed_buf(hipStream_t st, int kind, size_t need, Buffer** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return SAT_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  ManagedWs& m = g_ws[std::make_tuple(dev, (uintptr_t)st, kind)];
  if (!m.cur || m.cur->bytes < need) {
    size_t nb = std::max(need, (size_t)1 << 20);
    if (m.cur) nb = std::max(nb, 2 * m.cur->bytes);
    nb = (nb + 0xffff) & ~(size_t)0xffff;
    void* p = nullptr;
    if (hipMalloc(&p, nb) != hipSuccess) return SAT_ERR_HIP;
    if (kind == WS_CARRY && launch_clear(p, nb, st) != SAT_OK) {
      (void)hipFree(p);
      return SAT_ERR_HIP;
    }
    Buffer* b = new Buffer;
    b->ptr = p;
    b->bytes = nb;
    if (m.cur) m.retired.push_back(m.cur);
    m.cur = b;
  }
  *out = m.cur;
  return SAT_OK;
}

// Sticky give-up flags in pinned host memory, one word per device (see
// satdev::Fault); allocated on first use.
constexpr int MAX_DEVICES = 64;
static std::mutex g_hf_mu;
static uint32_t* g_hf_host = nullptr;  // host view
static uint32_t* g_hf_dev[MAX_DEVICES] = {};

uint32_t* host_fault_word(bool may_alloc) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_hf_mu);
  if (!g_hf_host) {
    if (!may_alloc) return nullptr;
    void* p = nullptr;
    if (hipHostMalloc(&p, MAX_DEVICES * sizeof(uint32_t),
                      hipHostMallocMapped | hipHostMallocPortable | hipHostMallocCoherent) != hipSuccess)
      return nullptr;
    std::memset(p, 0, MAX_DEVICES * sizeof(uint32_t));
    g_hf_host = (uint32_t*)p;
  }
  if (!g_hf_dev[dev]) {
    void* d = nullptr;
    if (hipHostGetDevicePointer(&d, g_hf_host, 0) != hipSuccess) return nullptr;
    g_hf_dev[dev] = (uint32_t*)d + dev;
  }
  return g_hf_dev[dev];
}

// Carry area of one API call and the epochs of its launches.
struct CarryCtx {
  void* ws = nullptr;
  size_t bytes = 0;        // usable bytes of the area (header + carries)
  Buffer* managed = nullptr;
  bool device_epoch = false;  // stream is being captured: bump on the device
  uint32_t local = 0;         // cleared caller area: next epoch
};

static bool stream_capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) return false;
  return cs == hipStreamCaptureStatusActive;
}

static sat_status carry_open(hipStream_t st, void* user_ws, size_t user_bytes, size_t need, CarryCtx* c) {
  if (user_ws) {
    if (user_bytes < need) return SAT_ERR_WORKSPACE;
    if ((uintptr_t)user_ws % 16) return SAT_ERR_INVALID;  // carry pairs are 16-byte accesses
    if (need > (size_t)satdev::WS_HEADER_BYTES) {
      const sat_status s = launch_clear((char*)user_ws + satdev::WS_HEADER_BYTES, need - satdev::WS_HEADER_BYTES, st);
      if (s != SAT_OK) return s;
    }
    c->ws = user_ws;
    c->bytes = need;
    c->local = 1;
    return SAT_OK;
  }
  const sat_status s = managed_buf(st, WS_CARRY, need, &c->managed);
  if (s != SAT_OK) return s;
  c->ws = c->managed->ptr;
  c->bytes = c->managed->bytes;
  c->device_epoch = stream_capturing(st);
  return SAT_OK;
}

// Epoch for the next launch on the call's carry area.
static sat_status carry_epoch(CarryCtx& c, hipStream_t st, uint32_t* ep) {
  if (c.local) {
    *ep = c.local++;
    return SAT_OK;
  }
  const uint64_t words = (c.bytes - satdev::WS_HEADER_BYTES) / 8;
  if (c.device_epoch) {
    *ep = 0;
    return launch_bump_epoch(c.ws, words, st);
  }
  std::lock_guard<std::mutex> lk(g_ws_mu);
  uint32_t e = ++c.managed->epoch;
  if (e >= satdev::EPOCH_DEVICE_BASE) {
    // host range exhausted: no stale tag may survive the restart at 1
    const sat_status s = launch_clear((char*)c.ws + satdev::WS_HEADER_BYTES, words * 8, st);
    if (s != SAT_OK) return s;
    e = c.managed->epoch = 1;
  }
  *ep = e;
  return SAT_OK;
}

// ---------------------------------------------------------------------------
// Validation helpers.
// ---------------------------------------------------------------------------
static bool pair_ok(int tin, int tout) { return pair_id(tin, tout) >= 0; }
static bool scan_ok(int scan) { return scan == SAT_SCAN_KOGGE_STONE || scan == SAT_SCAN_LF; }

static sat_status check_2d(const void* in, const void* out, int64_t ldi, int64_t ldo, uint32_t width,
                           uint32_t height, bool transposed_out) {
  if (width == 0 || height == 0) return SAT_OK;
  if (!in || !out) return SAT_ERR_NULL;
  if (ldi < (int64_t)width) return SAT_ERR_INVALID;
  if (ldo < (int64_t)(transposed_out ? height : width)) return SAT_ERR_INVALID;
  return SAT_OK;
}

static sat_status run_with_ws(Launch L, void* ws, size_t ws_bytes, hipStream_t st) {
  CarryCtx c;
  sat_status s = carry_open(st, ws, ws_bytes, carry_bytes(L), &c);
  if (s != SAT_OK) return s;
  s = carry_epoch(c, st, &L.epoch);
  if (s != SAT_OK) return s;
  return launch_family(L, c.ws, c.bytes, st);
}

static Launch make_launch(int family, int scan, int tin, int tout, const void* in, int64_t ldi, void* out, int64_t ldo,
                          uint32_t width, uint32_t height) {
  Launch L;
  L.family = family;
  L.scan = scan;
  L.tin = tin;
  L.tout = tout;
  L.in = in;
  L.out = out;
  L.ldi = ldi;
  L.ldo = ldo;
  L.in_plane = 0;
  L.out_plane = 0;
  L.width = width;
  L.height = height;
  L.planes = 1;
  L.epoch = 0;
  return L;
}

static sat_status pass_call(int family, int scan, int tin, int tout, const void* in, int64_t ldi, void* out,
                            int64_t ldo, uint32_t width, uint32_t height, void* stream) {
  if (!scan_ok(scan)) return SAT_ERR_INVALID;
  if (!pair_ok(tin, tout)) return SAT_ERR_DTYPE;
  const bool tr = family == FAM_ROWS_T || family == FAM_SERIAL_T;
  const sat_status c = check_2d(in, out, ldi, ldo, width, height, tr);
  if (c != SAT_OK || width == 0 || height == 0) return c;
  return run_with_ws(make_launch(family, scan, tin, tout, in, ldi, out, ldo, width, height), nullptr, 0,
                     (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// Full SAT: workspace layout = carries (max over the passes) | temporary.
// ---------------------------------------------------------------------------
struct SatPlan {
  Launch p1, p2;   // p2.family < 0 for the one-pass algorithm
  size_t carry;    // bytes of header + carries
  size_t tmp_off;  // byte offset of the temporary
  size_t total;    // workspace bytes
  int64_t tmp_pitch, tmp_plane;
};

static int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

static sat_status plan_integral(int alg, int scan, int tin, int tout, const void* in, int64_t ldi, int64_t in_plane,
                                void* out, int64_t ldo, int64_t out_plane, uint32_t width, uint32_t height,
                                uint32_t planes, SatPlan* P) {
  if (!scan_ok(scan)) return SAT_ERR_INVALID;
  if (!pair_ok(tin, tout)) return SAT_ERR_DTYPE;
  const size_t osz = dtype_size(tout);
  Launch a = make_launch(FAM_SAT, scan, tin, tout, in, ldi, out, ldo, width, height);
  a.in_plane = in_plane;
  a.out_plane = out_plane;
  a.planes = planes;
  Launch b = a;
  b.family = -1;
  P->tmp_pitch = 0;
  P->tmp_plane = 0;
  switch (alg) {
    case SAT_ALG_SINGLE_PASS:
      a.cfg = tune_pick(dtype_size(tin), osz, width, height, planes);
      break;
    case SAT_ALG_SCANROW_BRLT:
    case SAT_ALG_BRLT_SCANROW: {
      // pass 1: in (H x W) -> tmp (W x H); pass 2: tmp -> out (H x W)
      const int fam = alg == SAT_ALG_SCANROW_BRLT ? FAM_ROWS_T : FAM_SERIAL_T;
      P->tmp_pitch = round_up(height, 64);
      P->tmp_plane = P->tmp_pitch * width;
      a.family = fam;
      a.out = nullptr;
      a.ldo = P->tmp_pitch;
      a.out_plane = P->tmp_plane;
      b = make_launch(fam, scan, tout, tout, nullptr, P->tmp_pitch, out, ldo, height, width);
      b.in_plane = P->tmp_plane;
      b.out_plane = out_plane;
      b.planes = planes;
      break;
    }
    case SAT_ALG_SCANROW_COLUMN: {
      P->tmp_pitch = round_up(width, 64);
      P->tmp_plane = P->tmp_pitch * height;
      a.family = FAM_ROWS;
      a.out = nullptr;
      a.ldo = P->tmp_pitch;
      a.out_plane = P->tmp_plane;
      b = make_launch(FAM_COLS, scan, tout, tout, nullptr, P->tmp_pitch, out, ldo, width, height);
      b.in_plane = P->tmp_plane;
      b.out_plane = out_plane;
      b.planes = planes;
      break;
    }
    default:
      return SAT_ERR_INVALID;
  }
  P->p1 = a;
  P->p2 = b;
  size_t carry = carry_bytes_any_cfg(a);
  if (b.family >= 0) carry = std::max(carry, carry_bytes(b));
  P->carry = carry;
  P->tmp_off = (size_t)round_up((int64_t)carry, 256);
  P->total = P->tmp_off + (b.family >= 0 ? (size_t)P->tmp_plane * planes * osz : 0);
  return SAT_OK;
}

static sat_status run_integral(SatPlan& P, void* ws, size_t ws_bytes, hipStream_t st) {
  void* tmp = nullptr;
  CarryCtx c;
  if (ws) {
    if (ws_bytes < P.total) return SAT_ERR_WORKSPACE;
    tmp = (char*)ws + P.tmp_off;
  } else if (P.p2.family >= 0) {
    Buffer* b = nullptr;
    const sat_status s = managed_buf(st, WS_TMP, P.total - P.tmp_off, &b);
    if (s != SAT_OK) return s;
    tmp = b->ptr;
  }
  sat_status s = carry_open(st, ws, ws ? P.carry : 0, P.carry, &c);
  if (s != SAT_OK) return s;
  if ((s = carry_epoch(c, st, &P.p1.epoch)) != SAT_OK) return s;
  if (P.p2.family < 0) return launch_family(P.p1, c.ws, c.bytes, st);
  P.p1.out = tmp;
  P.p2.in = tmp;
  if ((s = launch_family(P.p1, c.ws, c.bytes, st)) != SAT_OK) return s;
  if ((s = carry_epoch(c, st, &P.p2.epoch)) != SAT_OK) return s;
  return launch_family(P.p2, c.ws, c.bytes, st);
}

// ---------------------------------------------------------------------------
// Batch plans. Layout in the workspace:
//   [0, 256)          launch-state header
//   [256, items)      carry words
//   items             BatchItem[count]
//   queue             uint32_t[2 * blocks]: (image, tile of the image) per workgroup
//
// Queue order: workgroup b runs on XCD b % nx and each XCD starts its
// workgroups in increasing b, so images are dealt whole to XCDs (largest
// first, each to the XCD with the fewest bytes so far): every look-back stays
// on one die. An XCD walks its images tile row by tile row, first row of every
// image, then the second rows, ...: a tile's upper neighbour was started a
// whole round of rows earlier and has usually finished, so the column
// look-back reads one predecessor. The XCDs' shares are interleaved (block
// k*nx + x is the k-th item of XCD x) and padded with empty items to equal
// length. Each image's tiles stay row-major and in increasing block order, so
// the queue is also safe for any other dealing of workgroups that starts them
// in increasing b.
//
// When whole images cannot load the XCDs evenly, the largest are not dealt
// (the XCDs holding none would idle: 4 f32 8192^2 images ran at 48 % of HBM
// peak): such images come first in the queue, their tiles row-major at
// consecutive blocks 0, 1, 2, ... over all XCDs, and the dealt images follow
// from the first whole round of blocks after them. A look-back of such an
// image reads tiles of other XCDs, which have smaller block numbers. Plans
// holding such images take the one-workgroup-per-item kernel: the persistent
// one only ever waits within one XCD's list.
// ---------------------------------------------------------------------------
constexpr uint32_t QUEUE_EMPTY = 0xffffffffu;

struct BatchPlanInfo {
  int out_size = 0;
  int cfg = 0;
  uint32_t total_tiles = 0;
  uint32_t blocks = 0;  // queue items (tiles + empty items), items_per_xcd * nx
  uint32_t nx = 1;
  uint32_t spread = 0;  // images laid over all XCDs
  uint64_t carry_words = 0;
  size_t items_off = 0, queue_off = 0, tickets_off = 0, total = 0;
  bool persistent = false;
  uint32_t grid = 0;  // workgroups of the persistent form
};
// Batch plans prepared from now on use the persistent kernel (1) or one
// workgroup per queue item (0, the default: the hardware's workgroup
// dispatch refills freed slots faster than ticket counters hand out items;
// 512 uneven images: 0.95 ms against 1.18 ms persistent).
static std::atomic<int> g_batch_persistent{0};
static std::mutex g_batch_mu;
static std::map<uintptr_t, BatchPlanInfo> g_batch;

// Deal images `ids` to XCDs (see the layout comment): per-XCD lists of image
// ids, balanced by bytes moved.
static std::vector<std::vector<uint32_t>> deal_images(const std::vector<uint64_t>& bytes,
                                                      std::vector<uint32_t> order, int nx) {
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return bytes[a] > bytes[b]; });
  std::vector<std::vector<uint32_t>> lists(nx);
  std::vector<uint64_t> load(nx, 0);
  for (uint32_t i : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lists[x].push_back(i);
    load[x] += bytes[i];
  }
  return lists;
}

static sat_status batch_layout(const sat_image* im, int count, int cfg, BatchPlanInfo* info,
                               std::vector<BatchItem>* items, std::vector<uint32_t>* queue) {
  if (count < 0) return SAT_ERR_INVALID;
  if (count > 0 && !im) return SAT_ERR_NULL;
  int osz = 0;
  for (int i = 0; i < count; ++i) {
    if (!pair_ok(im[i].tin, im[i].tout)) return SAT_ERR_DTYPE;
    const int s = (int)dtype_size(im[i].tout);
    if (osz == 0) osz = s;
    if (s != osz) return SAT_ERR_DTYPE;
    const sat_status c =
        check_2d(im[i].in, im[i].out, im[i].in_pitch, im[i].out_pitch, im[i].width, im[i].height, false);
    if (c != SAT_OK) return c;
  }
  if (osz == 0) osz = 4;
  if (cfg < 0) cfg = batch_cfg_pick(osz);
  const Geometry g = batch_geometry(osz, cfg);
  uint64_t tiles = 0;
  std::vector<uint64_t> ntiles(count);
  for (int i = 0; i < count; ++i) tiles += (ntiles[i] = tiles_of(g, im[i].width, im[i].height, 1));
  if (tiles > 0x7fffffffull) return SAT_ERR_INVALID;
  const int nx = std::max(1, device_xcd_count());
  std::vector<uint64_t> bytes(count);
  for (int i = 0; i < count; ++i)
    bytes[i] = (uint64_t)im[i].width * im[i].height * (dtype_size(im[i].tin) + dtype_size(im[i].tout));
  // The s largest images are spread over all XCDs and the rest dealt whole,
  // s the smallest count for which no XCD's dealt bytes exceed their mean by
  // more than 2 % (past 64 tries every image is spread).
  std::vector<uint32_t> by_size(count);
  for (int i = 0; i < count; ++i) by_size[i] = (uint32_t)i;
  std::stable_sort(by_size.begin(), by_size.end(), [&](uint32_t a, uint32_t b) { return bytes[a] > bytes[b]; });
  size_t ns = 0;
  std::vector<std::vector<uint32_t>> lists;
  for (;; ++ns) {
    if (nx > 1 && ns > 64) ns = (size_t)count;
    const std::vector<uint32_t> rest(by_size.begin() + ns, by_size.end());
    lists = deal_images(bytes, rest, nx);
    if (nx == 1 || ns == (size_t)count) break;
    uint64_t sum = 0, top = 0;
    for (int x = 0; x < nx; ++x) {
      uint64_t b = 0;
      for (uint32_t i : lists[x]) b += bytes[i];
      sum += b;
      top = std::max(top, b);
    }
    if (top * (uint64_t)nx * 50 <= sum * 51) break;
  }
  const std::vector<uint32_t> spread(by_size.begin(), by_size.begin() + ns);
  uint64_t spread_tiles = 0;
  for (uint32_t i : spread) spread_tiles += ntiles[i];
  const uint64_t k0 = (spread_tiles + nx - 1) / nx;  // rounds of blocks the spread images take
  uint64_t longest = 0;
  for (int x = 0; x < nx; ++x) {
    uint64_t n = 0;
    for (uint32_t i : lists[x]) n += ntiles[i];
    longest = std::max(longest, n);
  }
  longest += k0;
  const uint64_t blocks = longest * (uint64_t)nx;
  if (blocks > 0x7fffffffull) return SAT_ERR_INVALID;
  info->out_size = osz;
  info->cfg = cfg;
  info->total_tiles = (uint32_t)tiles;
  info->blocks = (uint32_t)blocks;
  info->nx = (uint32_t)nx;
  info->spread = (uint32_t)spread.size();
  info->carry_words = tiles * (uint64_t)(g.vwords + g.hwords);
  info->items_off = (size_t)round_up(satdev::WS_HEADER_BYTES + (int64_t)info->carry_words * 8, 256);
  info->queue_off = (size_t)round_up((int64_t)(info->items_off + sizeof(BatchItem) * (size_t)count), 256);
  info->tickets_off = (size_t)round_up((int64_t)(info->queue_off + 2 * sizeof(uint32_t) * (size_t)blocks), 256);
  info->total = info->tickets_off + (size_t)nx * satdev::BATCH_TICKET_STRIDE * sizeof(uint32_t);
  if (!items) return SAT_OK;
  items->resize(count);
  queue->assign(2 * blocks, QUEUE_EMPTY);
  uint64_t pos = 0;
  for (uint32_t i : spread)
    for (uint64_t t = 0; t < ntiles[i]; ++t, ++pos) {
      (*queue)[2 * pos] = i;
      (*queue)[2 * pos + 1] = (uint32_t)t;
    }
  for (int x = 0; x < nx; ++x) {
    uint64_t k = k0, rows = 0;
    for (uint32_t i : lists[x]) rows = std::max<uint64_t>(rows, (im[i].height + g.th - 1) / g.th);
    for (uint64_t r = 0; r < rows; ++r)
      for (uint32_t i : lists[x]) {
        const uint64_t nTx = (im[i].width + g.tw - 1) / g.tw;
        if (r >= ntiles[i] / nTx) continue;
        for (uint64_t c = 0; c < nTx; ++c, ++k) {
          (*queue)[2 * (k * nx + x)] = i;
          (*queue)[2 * (k * nx + x) + 1] = (uint32_t)(r * nTx + c);
        }
      }
  }
  uint32_t base = 0;
  for (int i = 0; i < count; ++i) {
    BatchItem& it = (*items)[i];
    const uint32_t nTx = (im[i].width + g.tw - 1) / g.tw, nTy = (im[i].height + g.th - 1) / g.th;
    it.g.in = im[i].in;
    it.g.out = im[i].out;
    it.g.ldi = im[i].in_pitch;
    it.g.ldo = im[i].out_pitch;
    it.g.H = im[i].height;
    it.g.W = im[i].width;
    it.g.nTx = nTx;
    it.g.nTy = nTy;
    it.g.vcarry = nullptr;  // rebased on the device address in sat_batch_prepare
    it.g.hcarry = nullptr;
    it.g.fault.ws = nullptr;  // set on the device address in sat_batch_prepare
    it.g.fault.host = nullptr;
    it.kind = (uint32_t)pair_id(im[i].tin, im[i].tout);
    it.tile_base = base;
    base += nTx * nTy;
  }
  return SAT_OK;
}

// ---------------------------------------------------------------------------
// Workspace of sat_integral_noalloc: a file-scope device array, so that the
// entry point allocates nothing. Its device address on each device is looked
// up once (hipGetSymbolAddress, with that device current) on the entry
// point's first call there; host code uses only that address. Every call on
// a device shares the device's array, which keeps the last call's contents.
// ---------------------------------------------------------------------------
constexpr size_t SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[MAX_DEVICES] = {};

// The current device's address of `scratch`; on the first call for a device
// its header (launch state and fault word) is zeroed on `st`, a fill kernel.
static sat_status scratch_on_device(hipStream_t st, void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return SAT_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr) return SAT_ERR_HIP;
    const sat_status s = launch_clear(ptr, satdev::WS_HEADER_BYTES, st);
    if (s != SAT_OK) return s;
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return SAT_OK;
}

}  // namespace sathost

using namespace sathost;

extern "C" {

int sat_abi_version(void) { return SAT_ABI_VERSION; }

size_t sat_dtype_size(int dtype) { return dtype_size(dtype); }

int sat_dtype_pair_supported(int tin, int tout) { return pair_id(tin, tout) >= 0 ? 1 : 0; }

const char* sat_status_string(int status) {
  switch (status) {
    case SAT_OK: return "ok";
    case SAT_ERR_NULL: return "null pointer";
    case SAT_ERR_INVALID: return "invalid argument";
    case SAT_ERR_DTYPE: return "unsupported dtype pair";
    case SAT_ERR_WORKSPACE: return "workspace too small";
    case SAT_ERR_HIP: return "HIP runtime error";
    case SAT_ERR_UNSUPPORTED: return "not supported by this build";
    case SAT_ERR_LOOKBACK: return "a look-back gave up waiting for a carry: results invalid";
    default: return "unknown status";
  }
}

int sat_tile_config_count(void) { return TILE_CFGS; }

void sat_tune_class(uint32_t width, uint32_t height, uint32_t planes, int* class_x, int* class_y, int* class_p) {
  int cx, cy, cp;
  tune_class(width, height, planes, &cx, &cy, &cp);
  if (class_x) *class_x = cx;
  if (class_y) *class_y = cy;
  if (class_p) *class_p = cp;
}

int sat_tune_get(int in_bytes, int out_bytes, int class_x, int class_y, int class_p) {
  if (!widths_ok(in_bytes, out_bytes) || !class_ok(class_x, class_y, class_p)) return -SAT_ERR_INVALID;
  return tune_lookup(in_bytes, out_bytes, class_x, class_y, class_p);
}

int sat_tune_set(int in_bytes, int out_bytes, int class_x, int class_y, int class_p, int cfg) {
  if (!widths_ok(in_bytes, out_bytes) || !class_ok(class_x, class_y, class_p) || cfg < -1 || cfg >= TILE_CFGS)
    return SAT_ERR_INVALID;
  g_tune[in_index(in_bytes)][out_bytes == 8 ? 1 : 0][class_x][class_y][class_p].store(cfg + 1,
                                                                                      std::memory_order_relaxed);
  return SAT_OK;
}

int sat_tune_pick(int in_bytes, int out_bytes, uint32_t width, uint32_t height, uint32_t planes) {
  if (!widths_ok(in_bytes, out_bytes)) return -SAT_ERR_INVALID;
  return tune_pick((size_t)in_bytes, (size_t)out_bytes, width, height, planes);
}

int sat_batch_tune_get(int out_bytes) {
  if (out_bytes != 4 && out_bytes != 8) return -SAT_ERR_INVALID;
  return batch_cfg_pick((size_t)out_bytes);
}

int sat_batch_tune_set(int out_bytes, int cfg) {
  if ((out_bytes != 4 && out_bytes != 8) || cfg < -1 || cfg >= BATCH_CFGS) return SAT_ERR_INVALID;
  g_tune_batch[out_bytes == 8 ? 1 : 0].store(cfg + 1, std::memory_order_relaxed);
  return SAT_OK;
}

static int read_clear_fault(void* workspace, hipStream_t st) {
  uint32_t v = 0;
  uint32_t* p = &((satdev::WsHeader*)workspace)->fault;
  if (hipStreamSynchronize(st) != hipSuccess) return -SAT_ERR_HIP;
  if (hipMemcpy(&v, p, sizeof v, hipMemcpyDeviceToHost) != hipSuccess) return -SAT_ERR_HIP;
  if (v) {
    const uint32_t z = 0;
    if (hipMemcpy(p, &z, sizeof z, hipMemcpyHostToDevice) != hipSuccess) return -SAT_ERR_HIP;
  }
  return v ? 1 : 0;
}

int sat_workspace_fault(void* workspace, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (workspace) return read_clear_fault(workspace, st);
  // the stream's managed carry buffers: the current one and any it replaced
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -SAT_ERR_HIP;
  std::vector<void*> bufs;
  {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto f = g_ws.find(std::make_tuple(dev, (uintptr_t)st, (int)WS_CARRY));
    if (f == g_ws.end() || !f->second.cur) return 0;
    bufs.push_back(f->second.cur->ptr);
    for (Buffer* b : f->second.retired) bufs.push_back(b->ptr);
  }
  int any = 0;
  for (void* b : bufs) {
    const int r = read_clear_fault(b, st);
    if (r < 0) return r;
    any |= r;
  }
  return any;
}

int sat_device_fault(int clear) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return -SAT_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_hf_mu);
  if (!g_hf_host) return 0;
  volatile uint32_t* w = g_hf_host + dev;
  const uint32_t v = *w;
  if (v && clear) *w = 0;
  return v ? 1 : 0;
}

int sat_release_workspaces(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return SAT_ERR_HIP;
  if (hipDeviceSynchronize() != hipSuccess) return SAT_ERR_HIP;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  for (auto it = g_ws.begin(); it != g_ws.end();) {
    if (std::get<0>(it->first) != dev) {
      ++it;
      continue;
    }
    ManagedWs& m = it->second;
    for (Buffer* b : m.retired) {
      (void)hipFree(b->ptr);
      delete b;
    }
    if (m.cur) {
      (void)hipFree(m.cur->ptr);
      delete m.cur;
    }
    it = g_ws.erase(it);
  }
  return SAT_OK;
}

int sat_stream_copy(const void* src, void* dst, size_t bytes, void* stream) {
  if (bytes == 0) return SAT_OK;
  if (!src || !dst) return SAT_ERR_NULL;
  if (((uintptr_t)src | (uintptr_t)dst) % 16 != 0) return SAT_ERR_INVALID;
  const uintptr_t s0 = (uintptr_t)src, d0 = (uintptr_t)dst;
  if (s0 < d0 + bytes && d0 < s0 + bytes) return SAT_ERR_INVALID;  // overlapping buffers
  return launch_stream_copy(src, dst, bytes, (hipStream_t)stream);
}

int sat_block_scan(int scan, int tin, int tout, const void* in, void* out, uint32_t width, uint32_t widthStride,
                   uint32_t height, uint32_t heightStride, void* stream) {
  return pass_call(FAM_ROWS_T, scan, tin, tout, in, widthStride, out, heightStride, width, height, stream);
}

int sat_seriel_scan(int tin, int tout, const void* in, void* out, uint32_t width, uint32_t widthStride,
                    uint32_t height, uint32_t heightStride, void* stream) {
  return pass_call(FAM_SERIAL_T, SAT_SCAN_KOGGE_STONE, tin, tout, in, widthStride, out, heightStride, width, height,
                   stream);
}

int sat_scan_x(int scan, int tin, int tout, const void* in, void* out, uint32_t width, uint32_t widthStride,
               uint32_t height, void* stream) {
  return pass_call(FAM_ROWS, scan, tin, tout, in, widthStride, out, widthStride, width, height, stream);
}

int sat_scan_y(int scan, int tin, int tout, const void* in, void* out, uint32_t width, uint32_t widthStride,
               uint32_t height, void* stream) {
  return pass_call(FAM_COLS, scan, tin, tout, in, widthStride, out, widthStride, width, height, stream);
}

int sat_scan_rows(int scan, int tin, int tout, const void* in, int64_t in_pitch, void* out, int64_t out_pitch,
                  uint32_t width, uint32_t height, void* stream) {
  return pass_call(FAM_ROWS, scan, tin, tout, in, in_pitch, out, out_pitch, width, height, stream);
}

int sat_scan_cols(int scan, int tin, int tout, const void* in, int64_t in_pitch, void* out, int64_t out_pitch,
                  uint32_t width, uint32_t height, void* stream) {
  return pass_call(FAM_COLS, scan, tin, tout, in, in_pitch, out, out_pitch, width, height, stream);
}

size_t sat_integral_planes_workspace_bytes(int alg, int tin, int tout, uint32_t width, uint32_t height,
                                           uint32_t planes) {
  SatPlan P;
  int dummy = 0;
  if (plan_integral(alg, SAT_SCAN_KOGGE_STONE, tin, tout, &dummy, width, (int64_t)width * height, &dummy, width,
                    (int64_t)width * height, width, height, planes, &P) != SAT_OK)
    return 0;
  return P.total;
}

size_t sat_integral_workspace_bytes(int alg, int tin, int tout, uint32_t width, uint32_t height) {
  return sat_integral_planes_workspace_bytes(alg, tin, tout, width, height, 1);
}

int sat_integral_planes(int alg, int scan, int tin, int tout, const void* in, int64_t in_pitch,
                        int64_t in_plane_stride, void* out, int64_t out_pitch, int64_t out_plane_stride,
                        uint32_t width, uint32_t height, uint32_t planes, void* workspace, size_t workspace_bytes,
                        void* stream) {
  if (!scan_ok(scan) || alg < 0 || alg > SAT_ALG_SINGLE_PASS) return SAT_ERR_INVALID;
  if (!pair_ok(tin, tout)) return SAT_ERR_DTYPE;
  if (width == 0 || height == 0 || planes == 0) return SAT_OK;
  sat_status c = check_2d(in, out, in_pitch, out_pitch, width, height, false);
  if (c != SAT_OK) return c;
  if (planes > 1 && (in_plane_stride < in_pitch * (int64_t)height || out_plane_stride < out_pitch * (int64_t)height))
    return SAT_ERR_INVALID;
  SatPlan P;
  c = plan_integral(alg, scan, tin, tout, in, in_pitch, in_plane_stride, out, out_pitch, out_plane_stride, width,
                    height, planes, &P);
  if (c != SAT_OK) return c;
  return run_integral(P, workspace, workspace_bytes, (hipStream_t)stream);
}

int sat_integral(int alg, int scan, int tin, int tout, const void* in, int64_t in_pitch, void* out, int64_t out_pitch,
                 uint32_t width, uint32_t height, void* workspace, size_t workspace_bytes, void* stream) {
  return sat_integral_planes(alg, scan, tin, tout, in, in_pitch, 0, out, out_pitch, 0, width, height, 1, workspace,
                             workspace_bytes, stream);
}

int sat_integral_bordered_planes(int alg, int scan, int tin, int tout, const void* in, int64_t in_pitch,
                                 int64_t in_plane_stride, void* out, int64_t out_pitch, int64_t out_plane_stride,
                                 uint32_t width, uint32_t height, uint32_t planes, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  if (!scan_ok(scan) || alg < 0 || alg > SAT_ALG_SINGLE_PASS) return SAT_ERR_INVALID;
  if (!pair_ok(tin, tout)) return SAT_ERR_DTYPE;
  if (planes == 0) return SAT_OK;
  if (!out) return SAT_ERR_NULL;
  if (out_pitch < (int64_t)width + 1) return SAT_ERR_INVALID;
  if (planes > 1 && out_plane_stride < out_pitch * ((int64_t)height + 1)) return SAT_ERR_INVALID;
  const size_t osz = dtype_size(tout);
  hipStream_t st = (hipStream_t)stream;
  // the table proper starts one row and one column in
  if (width > 0 && height > 0) {
    const int s = sat_integral_planes(alg, scan, tin, tout, in, in_pitch, in_plane_stride,
                                      (char*)out + (size_t)(out_pitch + 1) * osz, out_pitch, out_plane_stride, width,
                                      height, planes, workspace, workspace_bytes, stream);
    if (s != SAT_OK) return s;
  }
  return launch_border(osz, out, out_pitch, out_plane_stride, width, height, planes, st);
}

int sat_integral_bordered(int alg, int scan, int tin, int tout, const void* in, int64_t in_pitch, void* out,
                          int64_t out_pitch, uint32_t width, uint32_t height, void* workspace, size_t workspace_bytes,
                          void* stream) {
  return sat_integral_bordered_planes(alg, scan, tin, tout, in, in_pitch, 0, out, out_pitch, 0, width, height, 1,
                                      workspace, workspace_bytes, stream);
}

size_t sat_integral_reverse_planes_workspace_bytes(int dtype, uint32_t width, uint32_t height, uint32_t planes) {
  Launch L = make_launch(FAM_SAT_REV, 0, dtype, dtype, nullptr, width, nullptr, width, width, height);
  L.planes = planes;
  return carry_bytes_any_cfg(L);
}

size_t sat_integral_reverse_workspace_bytes(int dtype, uint32_t width, uint32_t height) {
  return sat_integral_reverse_planes_workspace_bytes(dtype, width, height, 1);
}

int sat_integral_reverse_planes(int scan, int dtype, const void* in, int64_t in_pitch, int64_t in_plane_stride,
                                void* out, int64_t out_pitch, int64_t out_plane_stride, uint32_t width,
                                uint32_t height, uint32_t planes, void* workspace, size_t workspace_bytes,
                                void* stream) {
  if (!scan_ok(scan)) return SAT_ERR_INVALID;
  if (!(dtype == SAT_F32 || dtype == SAT_F64 || dtype == SAT_U32 || dtype == SAT_I32)) return SAT_ERR_DTYPE;
  if (width == 0 || height == 0 || planes == 0) return SAT_OK;
  const sat_status c = check_2d(in, out, in_pitch, out_pitch, width, height, false);
  if (c != SAT_OK) return c;
  if (planes > 1 && (in_plane_stride < in_pitch * (int64_t)height || out_plane_stride < out_pitch * (int64_t)height))
    return SAT_ERR_INVALID;
  Launch L = make_launch(FAM_SAT_REV, scan, dtype, dtype, in, in_pitch, out, out_pitch, width, height);
  L.in_plane = in_plane_stride;
  L.out_plane = out_plane_stride;
  L.planes = planes;
  L.cfg = tune_pick(dtype_size(dtype), dtype_size(dtype), width, height, planes);
  return run_with_ws(L, workspace, workspace_bytes, (hipStream_t)stream);
}

int sat_integral_reverse(int scan, int dtype, const void* in, int64_t in_pitch, void* out, int64_t out_pitch,
                         uint32_t width, uint32_t height, void* workspace, size_t workspace_bytes, void* stream) {
  return sat_integral_reverse_planes(scan, dtype, in, in_pitch, 0, out, out_pitch, 0, width, height, 1, workspace,
                                     workspace_bytes, stream);
}

size_t sat_batch_workspace_bytes(const sat_image* images, int count) {
  // the largest layout over the tile configurations: valid whatever the
  // tuning table picks at sat_batch_prepare
  size_t m = 0;
  for (int c = 0; c < BATCH_CFGS; ++c) {
    BatchPlanInfo info;
    if (batch_layout(images, count, c, &info, nullptr, nullptr) != SAT_OK) return 0;
    m = std::max(m, info.total);
  }
  return m;
}

int sat_batch_prepare(const sat_image* images, int count, void* workspace, size_t workspace_bytes, void* stream) {
  if (!workspace) return SAT_ERR_NULL;
  BatchPlanInfo info;
  std::vector<BatchItem> items;
  std::vector<uint32_t> queue;
  sat_status s = batch_layout(images, count, -1, &info, &items, &queue);
  if (s != SAT_OK) return s;
  if (workspace_bytes < info.total) return SAT_ERR_WORKSPACE;
  if ((uintptr_t)workspace % 16) return SAT_ERR_INVALID;
  const Geometry g = batch_geometry(info.out_size, info.cfg);
  uint64_t* carries = (uint64_t*)((char*)workspace + satdev::WS_HEADER_BYTES);
  for (auto& it : items) {
    it.g.fault.ws = &((satdev::WsHeader*)workspace)->fault;
    it.g.fault.host = host_fault_word();
    it.g.vcarry = carries + (uint64_t)it.tile_base * g.vwords;
    it.g.hcarry = carries + (uint64_t)info.total_tiles * g.vwords + (uint64_t)it.tile_base * g.hwords;
  }
  hipStream_t st = (hipStream_t)stream;
  if (launch_clear(workspace, info.items_off, st) != SAT_OK) return SAT_ERR_HIP;
  if (!items.empty() &&
      hipMemcpyAsync((char*)workspace + info.items_off, items.data(), sizeof(BatchItem) * items.size(),
                     hipMemcpyHostToDevice, st) != hipSuccess)
    return SAT_ERR_HIP;
  if (!queue.empty() &&
      hipMemcpyAsync((char*)workspace + info.queue_off, queue.data(), sizeof(uint32_t) * queue.size(),
                     hipMemcpyHostToDevice, st) != hipSuccess)
    return SAT_ERR_HIP;
  if (hipStreamSynchronize(st) != hipSuccess) return SAT_ERR_HIP;
  info.persistent = g_batch_persistent.load(std::memory_order_relaxed) != 0 && info.nx <= 256 && info.spread == 0;
  if (info.persistent) {
    // as many workgroups per XCD as are resident at once (more would only
    // wait), never more than the XCD's queue items
    const int per_cu = batch_resident_per_cu(info.out_size, info.cfg);
    const uint64_t per_xcd = (uint64_t)std::max(1, per_cu) * std::max(1, device_cu_count() / (int)info.nx);
    info.grid = (uint32_t)(std::min<uint64_t>(per_xcd, info.blocks / info.nx) * info.nx);
    if (info.grid == 0) info.persistent = false;
  }
  std::lock_guard<std::mutex> lk(g_batch_mu);
  g_batch[(uintptr_t)workspace] = info;
  return SAT_OK;
}

int sat_batch_run(void* workspace, void* stream) {
  if (!workspace) return SAT_ERR_NULL;
  BatchPlanInfo info;
  {
    std::lock_guard<std::mutex> lk(g_batch_mu);
    auto f = g_batch.find((uintptr_t)workspace);
    if (f == g_batch.end()) return SAT_ERR_INVALID;
    info = f->second;
  }
  if (info.total_tiles == 0) return SAT_OK;
  hipStream_t st = (hipStream_t)stream;
  BatchPersist pers;
  if (info.persistent) {
    pers.tickets = (uint32_t*)((char*)workspace + info.tickets_off);
    pers.grid = info.grid;
    pers.items_per_xcd = info.blocks / info.nx;
    pers.nx = info.nx;
  }
  const sat_status s = launch_bump_epoch(workspace, info.carry_words, st, pers.tickets, pers.tickets ? info.nx : 0);
  if (s != SAT_OK) return s;
  return launch_batch(info.out_size, info.cfg, (const BatchItem*)((char*)workspace + info.items_off),
                      (const uint32_t*)((char*)workspace + info.queue_off), info.blocks, workspace, pers, st);
}

int sat_batch_persistent_set(int on) {
  g_batch_persistent.store(on ? 1 : 0, std::memory_order_relaxed);
  return SAT_OK;
}

int sat_batch_persistent_get(void) { return g_batch_persistent.load(std::memory_order_relaxed); }

int sat_box_sums(int dtype, const void* sat, int64_t pitch, int64_t plane_stride, uint32_t width, uint32_t height,
                 uint32_t planes, const int32_t* boxes, uint32_t nbox, void* out, void* stream) {
  if (dtype != SAT_F32 && dtype != SAT_F64) return SAT_ERR_DTYPE;
  if ((uint64_t)planes * nbox == 0) return SAT_OK;
  if (!sat || !boxes || !out) return SAT_ERR_NULL;
  if (width == 0 || height == 0 || pitch < (int64_t)width) return SAT_ERR_INVALID;
  if (planes > 1 && plane_stride < pitch * (int64_t)height) return SAT_ERR_INVALID;
  return launch_box_sums(dtype, sat, pitch, plane_stride, width, height, planes, boxes, nbox, out,
                         (hipStream_t)stream);
}

int sat_box_scatter(int dtype, const int32_t* boxes, uint32_t nbox, const void* grad_out, void* grad_sat,
                    int64_t pitch, int64_t plane_stride, uint32_t width, uint32_t height, uint32_t planes,
                    void* stream) {
  if (dtype != SAT_F32 && dtype != SAT_F64) return SAT_ERR_DTYPE;
  if ((uint64_t)planes * nbox == 0 || width == 0 || height == 0) return SAT_OK;
  if (!boxes || !grad_out || !grad_sat) return SAT_ERR_NULL;
  if (pitch < (int64_t)width) return SAT_ERR_INVALID;
  if (planes > 1 && plane_stride < pitch * (int64_t)height) return SAT_ERR_INVALID;
  return launch_box_scatter(dtype, boxes, nbox, grad_out, grad_sat, pitch, plane_stride, width, height, planes,
                            (hipStream_t)stream);
}

// Workspace of sat_box_grad: header and carries, then (256-aligned) the
// impulse bins, which never share the carry area.
size_t sat_box_grad_workspace_bytes(int dtype, uint32_t width, uint32_t height, uint32_t planes, uint32_t nbox) {
  if (dtype != SAT_F32 && dtype != SAT_F64) return 0;
  const size_t carry = box_grad_carry_bytes(dtype, width, height, planes);
  const size_t bins = box_grad_bin_bytes(dtype, width, height, planes, nbox);
  return bins ? (size_t)round_up((int64_t)carry, 256) + bins : carry;
}

int sat_box_grad(int dtype, const int32_t* boxes, uint32_t nbox, const void* grad_out, void* grad_x, int64_t pitch,
                 int64_t plane_stride, uint32_t width, uint32_t height, uint32_t planes, void* workspace,
                 size_t workspace_bytes, void* stream) {
  if (dtype != SAT_F32 && dtype != SAT_F64) return SAT_ERR_DTYPE;
  if (width == 0 || height == 0 || planes == 0) return SAT_OK;
  if (!grad_x || (nbox > 0 && (!boxes || !grad_out))) return SAT_ERR_NULL;
  if (pitch < (int64_t)width) return SAT_ERR_INVALID;
  if (planes > 1 && plane_stride < pitch * (int64_t)height) return SAT_ERR_INVALID;
  hipStream_t st = (hipStream_t)stream;
  const size_t need = box_grad_carry_bytes(dtype, width, height, planes);
  const size_t bin_bytes = box_grad_bin_bytes(dtype, width, height, planes, nbox);
  void* bins = nullptr;
  if (workspace) {
    if (bin_bytes) {
      const size_t at = (size_t)round_up((int64_t)need, 256);
      if (workspace_bytes < at + bin_bytes) return SAT_ERR_WORKSPACE;
      bins = (char*)workspace + at;
    }
  } else if (bin_bytes) {
    Buffer* b = nullptr;
    const sat_status s = managed_buf(st, WS_TMP, bin_bytes, &b);
    if (s != SAT_OK) return s;
    bins = b->ptr;
  }
  CarryCtx c;
  sat_status s = carry_open(st, workspace, workspace ? need : 0, need, &c);
  if (s != SAT_OK) return s;
  uint32_t ep = 0;
  if ((s = carry_epoch(c, st, &ep)) != SAT_OK) return s;
  return launch_box_grad(dtype, boxes, nbox, grad_out, grad_x, pitch, plane_stride, width, height, planes, c.ws,
                         c.bytes, bins, ep, st);
}

size_t sat_integral_noalloc_max_bytes(void) { return SCRATCH_BYTES; }

int sat_integral_noalloc(int scan, int tin, int tout, const void* in, int64_t in_pitch, void* out, int64_t out_pitch,
                         uint32_t width, uint32_t height, void* stream) {
  if (!scan_ok(scan)) return SAT_ERR_INVALID;
  if (!pair_ok(tin, tout)) return SAT_ERR_DTYPE;
  if (width == 0 || height == 0) return SAT_OK;
  sat_status c = check_2d(in, out, in_pitch, out_pitch, width, height, false);
  if (c != SAT_OK) return c;
  SatPlan P;
  c = plan_integral(SAT_ALG_SINGLE_PASS, scan, tin, tout, in, in_pitch, 0, out, out_pitch, 0, width, height, 1, &P);
  if (c != SAT_OK) return c;
  if (P.total > SCRATCH_BYTES) return SAT_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  void* ws = nullptr;
  if ((c = scratch_on_device(st, &ws)) != SAT_OK) return c;
  P.p1.may_alloc = false;
  return run_integral(P, ws, SCRATCH_BYTES, st);
}

}  // extern "C"
