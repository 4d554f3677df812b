/*
 * er_runtime.hip - the C-ABI (include/er_capi.h) over HIP: resource registries, device memory,
 * per-tracer frame state and pass sequencing.
 *
 * Registry semantics follow Cuda::List (Source/list.cuh:27-173): one map per kind, IDs from a
 * per-kind counter, a rebind with a known ID updates in place, an unbind of an unknown ID is
 * ignored, a 32nd live object is ignored. Device copies of the scene are built lazily: every
 * bind bumps a scene version, and the next pass of a tracer rebuilds its DevParams (the
 * reference instead re-uploads the ~21 KB tracer on every pass, core.cu:128). Lights and
 * textures are indexed by ID on the device, which equals the reference's list position while
 * nothing has been unbound.
 *
 * Every entry point is synchronous on return, as the reference's, and serialised by a mutex.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/er_capi.h"
#include "er_path.h"

using erd::DevParams;

extern "C" hipError_t erd_launch_trace(const DevParams* p, float4* frame, uint32_t* s1, uint32_t* s2, void* queue,
                                       unsigned int* counts, void* events, void* jobs, unsigned char* vis,
                                       unsigned long long* counters, int W, int H, int full, int count, int lut_lds,
                                       int split, size_t lds_bytes, size_t lds_bits, int shade_blocks,
                                       int primary_blocks, int shadow_blocks, int tile_shift, hipStream_t st,
                                       hipStream_t st2, hipEvent_t shaded, hipEvent_t after_primary,
                                       const uint32_t* order_in, uint32_t* cost, uint32_t* order_out, int heavy_cap,
                                       int lanes_per_ray, int shadow_lanes_per_ray, int heavy_lanes_per_ray);
extern "C" void erd_pass_record_sizes(size_t* scatter, size_t* event, size_t* job);
extern "C" hipError_t erd_launch_post(const DevParams* p, const float4* frame, float4* running, uchar4* display,
                                      int no_iterations, unsigned int* qcount, int W, int H, hipStream_t st);
extern "C" hipError_t erd_launch_macro_bits(const float* mc_sigma, const uint32_t* clipped, uint32_t* bits, int n, int words,
                                            hipStream_t st);
extern "C" hipError_t erd_launch_group_bits(uint32_t* bits, const uint32_t* unclipped, int mx, int my, int mz, int gx, int gy,
                                            int gz, int occ_base, int free_base, int words, hipStream_t st);
extern "C" hipError_t erd_launch_bilateral(const void* bp, const uchar4* in, uchar4* tmp, uchar4* out, int W, int H,
                                           hipStream_t st);
extern "C" size_t erd_bilateral_params_size(void);
extern "C" hipError_t erd_launch_build_cells(const uint16_t* vox, uint4* cells, int rx, int ry, int rz, hipStream_t st);
extern "C" hipError_t erd_launch_macro_minmax(const uint16_t* vox, uint32_t* mm, int rx, int ry, int rz, int mx, int my,
                                              int mz, int shift, hipStream_t st);
extern "C" hipError_t erd_launch_macro_sigma(const uint32_t* mm, const float* lut, int lut_max, float* out, int n,
                                             hipStream_t st);
extern "C" hipError_t erd_launch_gm_max(const DevParams* p, unsigned int* out, hipStream_t st);
extern "C" hipError_t erd_launch_batch_trace(const erd::BatchJob* J, int na, int full, int primary_blocks,
                                             int lds_words, int shade_blocks, hipStream_t stream,
                                             hipStream_t stream2, hipEvent_t shaded, int lpr, int shadow_lpr);
extern "C" hipError_t erd_launch_batch_post(const erd::BatchJob* J, int na, unsigned int round, int active_tiles,
                                            hipStream_t stream);
extern "C" hipError_t erd_launch_round_end(const erd::BatchJob* J, int njobs, hipStream_t stream);
extern "C" hipError_t erd_batch_params_upload(const erd::DevParams* host, int n);
extern "C" size_t erd_adj_partials_per_block(void);
extern "C" hipError_t erd_launch_primary_shade(const DevParams* p, float4* frame, uint32_t* s1, uint32_t* s2, void* queue,
                                               unsigned int* counts, void* events, void* jobs, int W, int H, int lut_lds,
                                               size_t lds_bytes, size_t lds_bits, int shade_blocks, int primary_blocks,
                                               int tile_shift, hipStream_t st);
extern "C" hipError_t erd_launch_adj_wavefront(const DevParams* p, unsigned int* counts, const void* jobs, const void* events,
                                               const void* queue, const float4* adj, float* tvis, float4* frame, uint2* list,
                                               unsigned int* lcount, const uint32_t* s1start, const uint32_t* s2start,
                                               float* partials, int blocks, int n, size_t lds_bytes, hipStream_t st);
extern "C" hipError_t erd_launch_adj_reset(unsigned int* counts, unsigned int* lcount, hipStream_t st);
extern "C" hipError_t erd_launch_autofocus(const DevParams* p, int u, int v, uint32_t s0, uint32_t s1, float* out,
                                           hipStream_t st);
extern "C" hipError_t erd_launch_adjoint(const DevParams* p, int mode, const float* lut_eval, const float* sigma_eval,
                                         const float4* adj, float4* value, float* partials, uint32_t* s1, uint32_t* s2,
                                         int W, int H, size_t lds_bytes, hipStream_t st);
extern "C" hipError_t erd_launch_adj_filter_t(const float* dldr, float4* out, int W, int H, float weight, hipStream_t st);
/* workgroups of 256 lanes in each adjoint kernel (grid-stride over jobs, events and pixels) */
#ifndef ER_ADJ_BLOCKS
#define ER_ADJ_BLOCKS 1024
#endif
extern "C" hipError_t erd_launch_adj_reduce(const float* partials, int nblocks, double* out, hipStream_t st);
extern "C" hipError_t erd_launch_zero_words(uint32_t* p, size_t words, hipStream_t st);

/* ---- device scratch of er_adjoint_lut_noalloc ----------------------------------------------
 * The no-allocation adjoint keeps every buffer beyond its inputs and outputs in one file-scope
 * device array: one pass's queues, the seeds, the adjoint image, the workgroup histograms and
 * the FP64 gradient sum. Its layout for a film of n pixels is below; the array is sized for
 * ER_NOALLOC_MAX_PIXELS (512 x 512), about 95 MiB, and each GPU the library runs on holds
 * its own copy until the process ends. */
#define ER_NOALLOC_MAX_PIXELS ((size_t)512 * 512)

struct NoAllocLayout {
    size_t d1, d2, s1s, s2s, dldr, adj, tvis, frame, queue, events, jobs, list, counts, lcount, partials, out, total;
};

static __host__ __device__ constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static __host__ __device__ constexpr NoAllocLayout noalloc_layout(size_t n)
{
    NoAllocLayout L{};
    size_t o = 0;
    L.d1 = o;       o += align256(n * sizeof(uint32_t));
    L.d2 = o;       o += align256(n * sizeof(uint32_t));
    L.s1s = o;      o += align256(n * sizeof(uint32_t));
    L.s2s = o;      o += align256(n * sizeof(uint32_t));
    L.dldr = o;     o += align256(n * 3 * sizeof(float));
    L.adj = o;      o += align256(n * sizeof(float4));
    L.tvis = o;     o += align256(2 * n * sizeof(float));
    L.frame = o;    o += align256(n * sizeof(float4));
    L.queue = o;    o += align256(n * sizeof(erd::ScatterRec));
    L.events = o;   o += align256(n * sizeof(erd::EventRec));
    L.jobs = o;     o += align256(2 * n * sizeof(erd::ShadowJob));
    L.list = o;     o += align256(n * sizeof(uint2));
    L.counts = o;   o += align256(4 * sizeof(uint32_t));
    L.lcount = o;   o += align256(sizeof(uint32_t));
    L.partials = o; o += align256((size_t)3 * ER_ADJ_BLOCKS * ER_TF_LUT_NODES * 4 * sizeof(float));
    L.out = o;      o += align256((size_t)ER_TF_LUT_NODES * 4 * sizeof(double));
    L.total = o;
    return L;
}

static __device__ __align__(256) unsigned char scratch[noalloc_layout(ER_NOALLOC_MAX_PIXELS).total];

namespace {

std::recursive_mutex g_mutex;
std::string g_last_error;

int set_error(int status, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

} /* namespace */

/* the error message of the library's other translation units (not exported) */
extern "C" __attribute__((visibility("hidden"))) int er_set_error_message(int status, const char* message)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    return set_error(status, "%s", message);
}

namespace {

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_error(ER_ERROR, "%s (%s)", hipGetErrorString(e_), #expr);             \
    } while (0)

/* Set while an entry point that must not allocate runs (er_adjoint_lut_noalloc): any device
 * buffer that would need a hipMalloc or hipFree then fails the call instead. */
bool g_no_alloc = false;
struct NoAllocScope {
    NoAllocScope() { g_no_alloc = true; }
    ~NoAllocScope() { g_no_alloc = false; }
};

template<class T> struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    int alloc(size_t count)
    {
        if (count == n && ptr) return ER_OK;
        if (g_no_alloc)
            return set_error(ER_ERROR, "this entry point allocates no device memory, but the tracer's device tables "
                                       "need (re)allocation: render one pass (or call er_adjoint_lut) after a change of "
                                       "volume, film or launch configuration first");
        release();
        if (count == 0) return ER_OK;
        HIP_TRY(hipMalloc((void**)&ptr, count * sizeof(T)));
        n = count;
        return ER_OK;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
};

/* ---- host-side restatements used to build device data (same op order as the kernels) -- */
float plf_eval_host(const er_plf& p, float x)   /* Source/plf.h:68-91 */
{
    const int n = p.count;
    if (n <= 0) return 0.0f;
    if (x < p.node_range[0]) return p.value[0];
    if (x > p.node_range[1]) return p.value[n - 1];
    for (int i = 1; i < n; ++i) {
        const float P1 = p.position[i - 1], P2 = p.position[i];
        const float LerpT = (x - P1) / (P2 - P1);
        if (x >= P1 && x < P2) return p.value[i - 1] + LerpT * (p.value[i] - p.value[i - 1]);
    }
    return 0.0f;
}

float shape_area(const er_shape& s)   /* Shape::Update (Source/shape.h:91-103) */
{
    const float PI_F = 3.141592654f;
    switch (s.type) {
        case ER_SHAPE_PLANE:    return s.size[0] * s.size[1];
        case ER_SHAPE_DISK:     return PI_F * (s.outer_radius * s.outer_radius);
        case ER_SHAPE_RING:     return PI_F * (s.outer_radius * s.outer_radius) - PI_F * (s.inner_radius * s.inner_radius);
        case ER_SHAPE_BOX:      return (2.0f * s.size[0] * s.size[1]) + (2.0f * s.size[0] * s.size[2]) + (2.0f * s.size[1] * s.size[2]);
        case ER_SHAPE_SPHERE:   return 4.0f * PI_F * (s.outer_radius * s.outer_radius);
        case ER_SHAPE_CYLINDER: return (2.0f * PI_F * (s.outer_radius * s.outer_radius)) + (2.0f * PI_F * s.outer_radius * s.size[2]);
    }
    return s.area;
}

/* ---- the -log(u) table: u = m 2^-23, value = -log(u) rounded once to float ------------- */
std::vector<float> g_neglog_host;
float* g_neglog_dev = nullptr;

void build_neglog_host()
{
    if (!g_neglog_host.empty()) return;
    const size_t n = (size_t)1 << ER_NEGLOG_BITS;
    g_neglog_host.resize(n);
    g_neglog_host[0] = INFINITY;
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([t, nt]() {
            for (size_t m = 1 + t; m < n; m += nt)
                g_neglog_host[m] = (float)(-logl(ldexpl((long double)m, -ER_NEGLOG_BITS)));
        });
    for (auto& x : th) x.join();
}

int ensure_neglog_device()
{
    if (g_neglog_dev) return ER_OK;
    if (g_no_alloc)
        return set_error(ER_ERROR, "this entry point allocates no device memory, but the library's -log table is not on "
                                   "the device yet: call er_init or render one pass first");
    build_neglog_host();
    const size_t bytes = g_neglog_host.size() * sizeof(float);
    HIP_TRY(hipMalloc((void**)&g_neglog_dev, bytes));
    HIP_TRY(hipMemcpy(g_neglog_dev, g_neglog_host.data(), bytes, hipMemcpyHostToDevice));
    return ER_OK;
}

/* ---- resources --------------------------------------------------------------------------- */
uint64_t g_scene_version = 1;
hipStream_t g_batch_stream = nullptr, g_batch_stream2 = nullptr;

/* Instantiated batch graphs, keyed by everything a graph bakes in: the jobs' buffers and
 * parameter blocks (device pointers), film and tile layout, pass counts, the launch
 * configuration and the estimator mode. A later batch of the same jobs reuses the graph after
 * uploading its job tables (pass counters) and parameter blocks again. */
struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf<erd::BatchJob> dJ[2];
    void release()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        dJ[0].release();
        dJ[1].release();
    }
};
std::map<std::string, BatchGraph> g_batch_graphs;
const size_t ER_BATCH_GRAPH_CACHE = 4;

void clear_batch_graphs()
{
    for (auto& kv : g_batch_graphs) kv.second.release();
    g_batch_graphs.clear();
}


/* Macrocells (bricks) of 2^shift cells a side: the min/max intensity of each, from which a
 * tracer derives its occupancy bits. A volume keeps one grid per brick size in use. */
struct MacroGrid {
    int mc[3] = {0, 0, 0};
    DevBuf<uint32_t> mm;
};

struct Volume {
    er_volume_desc desc{};
    DevBuf<uint16_t> vox;
    DevBuf<uint4> cells;
    std::map<int, MacroGrid> grids;
    int mc_shift = 3;               /* default brick: 8 cells, coarser for very large volumes */
    uint16_t vmax = 0;
    uint64_t upload = 0;            /* upload serial number (autotuner cache key) */
    ~Volume()
    {
        vox.release(); cells.release();
        for (auto& g : grids) g.second.mm.release();
    }
};

/* The occupancy bits (one per macrocell), the empty-space radii (4 bits per group of
 * 2^ER_GROUP_SHIFT macrocells a side), the groups' clip-free bits and, when it takes at most
 * 32 KB, the opacity table (one float per intensity up to the volume's maximum) are staged in
 * LDS by the march kernels; together they must fit 52.5 KB (105 LDS allocation blocks of 512
 * bytes, with one more for a kernel's own LDS), so that three march workgroups (the primary
 * kernel's register-bound occupancy) share a CU's 160 KB. */
const size_t ER_MAX_TABLES_BYTES = 53760;
bool brick_fits(const er_volume_desc& d, uint16_t vmax, int shift)
{
    const size_t E = (size_t)1 << shift, G = (size_t)1 << ER_GROUP_SHIFT;
    size_t mc[3];
    for (int i = 0; i < 3; ++i) mc[i] = ((size_t)std::max(d.resolution[i], 1) + E - 1) / E;
    const size_t n = mc[0] * mc[1] * mc[2];
    const size_t g = ((mc[0] + G - 1) / G) * ((mc[1] + G - 1) / G) * ((mc[2] + G - 1) / G);
    const size_t lut = ((size_t)vmax + 1) * sizeof(float);
    /* macrocell bits, group radii and group clip-free bits, as build_params lays them out */
    return 4 * ((n + 31) / 32 + (g + 7) / 8 + (g + 31) / 32) + (lut <= 32768 ? lut : 0) <= ER_MAX_TABLES_BYTES;
}

/* The state of one pass in flight: its frame estimate and its work queues. A tracer keeps
 * two, so that a pass's shadow marches, combine and post kernels (second stream) overlap the
 * next pass's primary and shading kernels (first stream). */
struct PassSet {
    DevBuf<float4> frame;
    DevBuf<uint32_t> qcount;        /* [0] scatter queue, [1] shadow jobs, [2] primary tiles */
    DevBuf<char> queue, events, jobs;
    DevBuf<unsigned char> vis;
    hipEvent_t shaded = nullptr;    /* first stream: primary + shading of this set's pass done */
    hipEvent_t posted = nullptr;    /* second stream: the pass's post done, the set is free */
    /* markers recorded where a timed pass records its timing events (nobody waits on them):
     * without them the passes ran 8% (CT) to 15% (large) slower than timed ones */
    hipEvent_t marks[4] = {nullptr, nullptr, nullptr, nullptr};
    bool pending = false;           /* posted has been recorded */
    void release()
    {
        frame.release(); qcount.release(); queue.release(); events.release(); jobs.release(); vis.release();
        if (shaded) (void)hipEventDestroy(shaded);
        if (posted) (void)hipEventDestroy(posted);
        for (hipEvent_t& m : marks) {
            if (m) (void)hipEventDestroy(m);
            m = nullptr;
        }
        shaded = posted = nullptr;
    }
};

struct Tracer {
    er_tracer_desc desc{};
    int W = 0, H = 0;
    PassSet set[2];
    int cur = 0;                    /* set of the last pass */
    DevBuf<float4> running;
    DevBuf<uchar4> display, disp_tmp, disp_filt;
    DevBuf<uint32_t> seeds1, seeds2;
    DevBuf<float> lut, mc_sigma, wc_sigma;
    DevBuf<uint32_t> mc_bits, clip_bits;
    size_t lds_bytes = 0;
    int lut_lds = 0;
    DevBuf<DevParams> params;
    DevBuf<unsigned long long> counters;
    DevBuf<unsigned int> gm_tmp;
    DevBuf<char> bilateral;
    DevParams host_params{};
    uint64_t built_version = 0;
    int built_volume = -1;
    float gm_max = 0.0f;
    int counting = 0;
    std::vector<float> tf_lut;       /* node table (er_set_tf_lut), empty when the PLFs are used */
    float tf_scale = 0.0f;
    DevBuf<float> tf_lut_dev;
    /* launch configuration (er_autotune / er_set_tune): brick edge 2^mc_shift (0: the
     * volume's default), primary tiles 2^tile_shift pixels wide, persistent primary
     * workgroups per CU (0: the library default) */
    int mc_shift = 0, tile_shift = 3, blocks_per_cu = 0, shadow_waves_per_cu = 4, heavy_tiles = 0, rays_per_wave = 64,
        shadow_rays_per_wave = 64, heavy_rays_per_wave = 4;
    /* primary tiles: each pass records its tiles' march steps and orders the next pass's tiles
     * most expensive first; the order is valid for the tile shape and count it was built for */
    DevBuf<uint32_t> tile_cost, tile_order;
    int order_shift = -1, order_tiles = 0;
    hipStream_t stream = nullptr;     /* primary + shading, uploads, readbacks */
    hipStream_t stream2 = nullptr;    /* shadow marches, combine, post */
    /* the timing events of er_render_passes and of the tuner's trials (four a pass), kept from
     * call to call so that creating them is not part of a timed render */
    std::vector<hipEvent_t> timing_events;
    ~Tracer()
    {
        for (hipEvent_t e : timing_events) (void)hipEventDestroy(e);
        set[0].release(); set[1].release();
        running.release(); display.release(); disp_tmp.release(); disp_filt.release();
        seeds1.release(); seeds2.release(); lut.release(); mc_sigma.release(); wc_sigma.release(); params.release();
        mc_bits.release(); clip_bits.release(); counters.release(); gm_tmp.release(); bilateral.release();
        tile_cost.release(); tile_order.release();
        if (stream) (void)hipStreamDestroy(stream);
        if (stream2) (void)hipStreamDestroy(stream2);
    }
};

/* at least n timing events in t.timing_events */
int timing_events(Tracer& t, size_t n)
{
    while (t.timing_events.size() < n) {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreate(&e));
        t.timing_events.push_back(e);
    }
    return ER_OK;
}

/* both streams of a tracer idle */
int sync_tracer(Tracer& t)
{
    HIP_TRY(hipStreamSynchronize(t.stream));
    if (t.stream2) HIP_TRY(hipStreamSynchronize(t.stream2));
    return ER_OK;
}

template<class T> struct Registry {
    std::map<int, T*> map;
    int counter = 0;
    T* find(int id)
    {
        if (id < 0) return nullptr;
        auto it = map.find(id);
        return it == map.end() ? nullptr : it->second;
    }
    void clear()
    {
        for (auto& kv : map) delete kv.second;
        map.clear();
        counter = 0;
    }
};

struct LightObj { er_light_desc desc; };
struct TextureObj { er_texture_desc desc; };
struct ClipObj { er_clipping_desc desc; };
struct ObjectObj { er_object_desc desc; };
struct BitmapObj {
    er_bitmap_desc desc;
    std::vector<uint8_t> pixels;
    DevBuf<uchar4> dev;          /* device copy, refreshed when a bind brings new pixels */
    bool dev_stale = true;
    ~BitmapObj() { dev.release(); }
};

Registry<Tracer> g_tracers;
Registry<Volume> g_volumes;
Registry<LightObj> g_lights;
Registry<TextureObj> g_textures;
Registry<ClipObj> g_clips;
Registry<ObjectObj> g_objects;
Registry<BitmapObj> g_bitmaps;
bool g_device_ready = false;
int g_list_capacity = ER_LIST_CAPACITY;   /* list.cuh:27 MaxSize; er_set_list_capacity */
int g_primary_blocks = 2048;   /* persistent primary workgroups: 7 per CU (LDS and VGPR bound) */
int g_cus = 256;

/* Device address of `scratch` per device. In host code the array's name is only a stand-in
 * for the symbol; the address each device's copy lives at comes from hipGetSymbolAddress, with
 * that device current, and only that address is passed to kernels and copies. */
const int ER_MAX_DEVICES = 64;
unsigned char* g_scratch_dev[ER_MAX_DEVICES] = {};

int scratch_for_current_device(unsigned char*& out)
{
    out = nullptr;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= ER_MAX_DEVICES) return set_error(ER_ERROR, "device %d is beyond the scratch table", dev);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(ER_ERROR, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[dev] = (unsigned char*)p;
    }
    out = g_scratch_dev[dev];
    return ER_OK;
}

int ensure_device()
{
    if (g_device_ready) return ER_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return set_error(ER_ERROR, "no HIP device available (%s)", hipGetErrorString(e));
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
    {
        g_cus = cus;
        g_primary_blocks = cus * 7;
    }
    /* set-up of the no-allocation entry point's scratch on this device */
    unsigned char* sp = nullptr;
    int st = scratch_for_current_device(sp);
    if (st) return st;
    g_device_ready = true;
    return ER_OK;
}

/* SplitMix64 of (seed ^ pixel) split in two non-zero 32-bit halves */
void seed_pixels(uint64_t seed, size_t n, std::vector<uint32_t>& s1, std::vector<uint32_t>& s2)
{
    s1.resize(n); s2.resize(n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t z = (seed ^ (uint64_t)i) + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        uint32_t a = (uint32_t)z, b = (uint32_t)(z >> 32);
        s1[i] = a ? a : 1u;
        s2[i] = b ? b : 1u;
    }
}

int tracer_seed(Tracer& t, uint64_t seed)
{
    std::vector<uint32_t> s1, s2;
    const size_t n = (size_t)t.W * t.H;
    seed_pixels(seed, n, s1, s2);
    HIP_TRY(hipMemcpy(t.seeds1.ptr, s1.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t.seeds2.ptr, s2.data(), n * 4, hipMemcpyHostToDevice));
    return ER_OK;
}

/* FrameBuffer::Resize (Source/framebuffer.h:45-68): buffers, zeroed estimates, new seeds */
int tracer_resize(Tracer& t, int W, int H, int id)
{
    if (W == t.W && H == t.H && t.set[0].frame.ptr) return ER_OK;
    if (W < 0 || H < 0) return set_error(ER_ERROR, "invalid film size %d x %d", W, H);
    int st;
    if ((st = sync_tracer(t))) return st;
    const size_t n = (size_t)W * H;
    if ((st = t.running.alloc(n)) || (st = t.display.alloc(n)) || (st = t.disp_tmp.alloc(n)) ||
        (st = t.disp_filt.alloc(n)) || (st = t.seeds1.alloc(n)) || (st = t.seeds2.alloc(n)))
        return st;
    size_t rs = 0, re = 0, rj = 0;
    erd_pass_record_sizes(&rs, &re, &rj);
    for (PassSet& S : t.set) {
        if ((st = S.frame.alloc(n)) || (st = S.queue.alloc(n * rs)) || (st = S.events.alloc(n * re)) ||
            (st = S.jobs.alloc(2 * n * rj)) || (st = S.vis.alloc(2 * n)) || (st = S.qcount.alloc(4)))
            return st;
        if (!S.shaded) HIP_TRY(hipEventCreateWithFlags(&S.shaded, hipEventDisableTiming));
        if (!S.posted) HIP_TRY(hipEventCreateWithFlags(&S.posted, hipEventDisableTiming));
        for (hipEvent_t& m : S.marks)
            if (!m) HIP_TRY(hipEventCreateWithFlags(&m, hipEventDisableTiming));
        S.pending = false;
        if (n) {
            HIP_TRY(hipMemset(S.frame.ptr, 0, n * sizeof(float4)));
            HIP_TRY(hipMemset(S.qcount.ptr, 0, 4 * sizeof(uint32_t)));
        }
    }
    t.cur = 0;
    t.W = W; t.H = H;
    if (n) {
        HIP_TRY(hipMemset(t.running.ptr, 0, n * sizeof(float4)));
        HIP_TRY(hipMemset(t.display.ptr, 0, n * sizeof(uchar4)));
        HIP_TRY(hipMemset(t.disp_filt.ptr, 0, n * sizeof(uchar4)));
        return tracer_seed(t, 0x45787052ull + (uint64_t)id);
    }
    return ER_OK;
}

void mat34(float dst[3][4], const float src[4][4])
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j) dst[i][j] = src[i][j];
}

float norm_len(const float v[3]) { return sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* Camera::Update (Source/camera.h:61-96) */
void camera_update(DevParams& P, const er_camera& c)
{
    float N[3], U[3], V[3], d[3];
    for (int i = 0; i < 3; ++i) d[i] = c.target[i] - c.pos[i];
    const float Ln = norm_len(d);
    for (int i = 0; i < 3; ++i) N[i] = d[i] / Ln;
    float cu[3] = {(N[1] * c.up[2]) - (N[2] * c.up[1]), (N[2] * c.up[0]) - (N[0] * c.up[2]), (N[0] * c.up[1]) - (N[1] * c.up[0])};
    const float Lu = norm_len(cu);
    for (int i = 0; i < 3; ++i) U[i] = cu[i] / Lu;
    float cv[3] = {(N[1] * U[2]) - (N[2] * U[1]), (N[2] * U[0]) - (N[0] * U[2]), (N[0] * U[1]) - (N[1] * U[0])};
    const float Lv = norm_len(cv);
    for (int i = 0; i < 3; ++i) V[i] = cv[i] / Lv;
    for (int i = 0; i < 3; ++i) { P.cam_pos[i] = c.pos[i]; P.cam_N[i] = N[i]; P.cam_U[i] = U[i]; P.cam_V[i] = V[i]; }
    P.focal = c.focal_distance == -1.0f ? Ln : c.focal_distance;
    const float Scale = tanf((0.5f * c.fov / 57.29577951308232f));
    const float Aspect = (float)c.film_size[1] / (float)c.film_size[0];
    float s00, s01, s10, s11;
    if (Aspect > 1.0f) { s00 = -Scale; s01 = Scale; s10 = -Scale * Aspect; s11 = Scale * Aspect; }
    else { s00 = -Scale / Aspect; s01 = Scale / Aspect; s10 = -Scale; s11 = Scale; }
    P.screen00 = s00;
    P.screen10 = s10;
    P.invscreen0 = (s01 - s00) / (float)c.film_size[0];
    P.invscreen1 = (s11 - s10) / (float)c.film_size[1];
    P.aperture = c.aperture_size;
    P.clip_near = c.clip_near;
    P.clip_far = c.clip_far;
    P.exposure = c.exposure;
    P.film_w = c.film_size[0];
    P.film_h = c.film_size[1];
}

int upload_volume(Volume& v)
{
    const er_volume_desc& d = v.desc;
    const int rx = d.resolution[0], ry = d.resolution[1], rz = d.resolution[2];
    const size_t n = (size_t)rx * ry * rz;
    if (rx <= 0 || ry <= 0 || rz <= 0) return set_error(ER_ERROR, "invalid volume resolution %d x %d x %d", rx, ry, rz);
    /* the cell records are indexed with 24-bit factors and a 32-bit index (er_device.h) */
    if ((size_t)(rx + 1) * (ry + 1) >= ((size_t)1 << 24) || (size_t)(rz + 1) >= ((size_t)1 << 24) ||
        (size_t)(rx + 1) * (ry + 1) * (rz + 1) >= ((size_t)1 << 32))
        return set_error(ER_ERROR, "volume %d x %d x %d too large: (Rx+1)(Ry+1) and Rz+1 must stay below 2^24 and the "
                                   "cell count below 2^32", rx, ry, rz);
    if (!d.voxels) return set_error(ER_ERROR, "volume voxels pointer is NULL");
    int st;
    if ((st = v.vox.alloc(n))) return st;
    HIP_TRY(hipMemcpy(v.vox.ptr, d.voxels, n * sizeof(uint16_t), hipMemcpyHostToDevice));
    uint16_t vmax = 0;
    for (size_t i = 0; i < n; ++i) vmax = std::max(vmax, d.voxels[i]);
    v.vmax = vmax;
    const size_t nc = (size_t)(rx + 1) * (ry + 1) * (rz + 1);
    if ((st = v.cells.alloc(nc))) return st;
    HIP_TRY(erd_launch_build_cells(v.vox.ptr, v.cells.ptr, rx, ry, rz, 0));
    /* default macrocell edge: 8 cells, coarser when the occupancy bits and the opacity table
     * would not fit their 52.5 KB of LDS */
    v.mc_shift = 3;
    while (!brick_fits(d, v.vmax, v.mc_shift) && v.mc_shift < 8) v.mc_shift++;
    static uint64_t uploads = 0;
    v.upload = ++uploads;
    for (auto& g : v.grids) g.second.mm.release();
    v.grids.clear();
    HIP_TRY(hipDeviceSynchronize());
    return ER_OK;
}

/* the volume's macrocell grid of 2^shift cells a side, built on first use */
MacroGrid* volume_grid(Volume& v, int shift, int& st)
{
    st = ER_OK;
    auto it = v.grids.find(shift);
    if (it != v.grids.end()) return &it->second;
    const er_volume_desc& d = v.desc;
    const int E = 1 << shift;
    MacroGrid& g = v.grids[shift];
    for (int i = 0; i < 3; ++i) g.mc[i] = (d.resolution[i] + E - 1) / E;
    if ((st = g.mm.alloc((size_t)g.mc[0] * g.mc[1] * g.mc[2]))) { v.grids.erase(shift); return nullptr; }
    hipError_t e = erd_launch_macro_minmax(v.vox.ptr, g.mm.ptr, d.resolution[0], d.resolution[1], d.resolution[2], g.mc[0],
                                           g.mc[1], g.mc[2], shift, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        g.mm.release();
        v.grids.erase(shift);
        st = set_error(ER_ERROR, "%s (macrocell grid)", hipGetErrorString(e));
        return nullptr;
    }
    return &g;
}

bool needs_gm(const er_tracer_desc& d)
{
    const int t = d.settings.shading_type;
    return d.settings.estimator == ER_ESTIMATOR_FULL && t != ER_SHADING_BRDF_ONLY && t != ER_SHADING_PHASE_ONLY;
}

/* A bound on the magnitudes in the ray arithmetic of every march: the rays start at the camera
 * (within the aperture), at scatter points inside the volume or on a surface object, and end
 * inside the volume's box. The rounding of a sample position against that box stays far below
 * 1e-5 of it; the clip-aware occupancy bits widen the edge macrocells by that much. */
double ray_origin_bound(const erd::DevParams& P, const std::vector<const er_object_desc*>& objs)
{
    double box = 0.0, diag2 = 0.0;
    for (int i = 0; i < 3; ++i) {
        box = std::max(box, std::max(std::fabs((double)P.minp[i]), std::fabs((double)P.maxp[i])));
        diag2 += ((double)P.maxp[i] - P.minp[i]) * ((double)P.maxp[i] - P.minp[i]);
    }
    double origin = std::sqrt((double)P.cam_pos[0] * P.cam_pos[0] + (double)P.cam_pos[1] * P.cam_pos[1] +
                              (double)P.cam_pos[2] * P.cam_pos[2]) + std::fabs((double)P.aperture);
    for (const er_object_desc* o : objs) {
        const er_shape& S = o->shape;
        const double ext = std::max({std::fabs((double)S.size[0]), std::fabs((double)S.size[1]),
                                     std::fabs((double)S.size[2]), std::fabs((double)S.outer_radius)});
        double b = 0.0;
        for (int i = 0; i < 3; ++i) {
            double lin = 0.0;
            for (int j = 0; j < 3; ++j) lin += std::fabs((double)S.tm[i][j]);
            b += std::fabs((double)S.tm[i][3]) + lin * ext;
        }
        origin = std::max(origin, b);
    }
    const double r = 2.0 * origin + 4.0 * box + std::sqrt(diag2);
    return r == r ? r : 1.0e300;
}

/* The clipping class of macrocell m: bit 0 (ER_MC_CLIPPED) when every sample the march can
 * assign to m is clipped away (clipped() holds for it), so the macrocell can be skipped like an
 * empty one; bit 1 (ER_MC_UNCLIPPED) when no such sample is clipped away, so the march need not
 * test it. A sample is assigned to m when its lattice coordinates, computed in float, fall in
 * m's lattice box; the box is widened by 0.01 lattice units against that rounding, and an edge
 * macrocell (which also takes the samples the clamp maps to it) extends one lattice unit plus
 * edge_ext world units beyond the volume: samples lie inside the volume's box up to the
 * rounding of the ray arithmetic, which edge_ext bounds. The test runs on the host, in double
 * precision, on the 8 corners of that world box. Per clipping object, the box lies inside the
 * (convex) shape when all 8 corners do, and outside it when it is separated from the shape (a
 * face plane of the box or cylinder, or the corners' bounding box beyond the sphere's or
 * cylinder's radius); every comparison keeps a margin E = 1e-6 x (sum of the magnitudes of the
 * terms) over the float rounding of the local coordinates in clipped(), so the float test of
 * every sample in the box agrees. A sample is clipped when it is inside a plain object or
 * outside an inverted one (clipped()); so m is wholly clipped when some object clips the whole
 * box and wholly unclipped when every object clips none of it. NaN anywhere leaves both bits
 * clear. */
#define ER_MC_CLIPPED 1
#define ER_MC_UNCLIPPED 2
int mc_clip_class(const DevParams& p, int m, double edge_ext)
{
    const int sh = p.mc_shift;
    const int ix[3] = {m % p.mc_dim[0], (m / p.mc_dim[0]) % p.mc_dim[1], m / (p.mc_dim[0] * p.mc_dim[1])};
    double lo[3], hi[3], amax[3];
    for (int a = 0; a < 3; ++a) {
        const double cw = 1.0 / ((double)p.invsize[a] * (double)p.resf[a]);   /* world size of one lattice unit */
        const double l0 = ix[a] == 0 ? -1.0 : (double)(ix[a] << sh) - 0.01;
        const double l1 = ix[a] == p.mc_dim[a] - 1 ? (double)p.resf[a] + 1.0 : (double)((ix[a] + 1) << sh) + 0.01;
        lo[a] = (double)p.minp[a] + l0 * cw - (ix[a] == 0 ? edge_ext : 0.0);
        hi[a] = (double)p.minp[a] + l1 * cw + (ix[a] == p.mc_dim[a] - 1 ? edge_ext : 0.0);
        amax[a] = std::fmax(std::fabs(lo[a]), std::fabs(hi[a]));
    }
    bool unclipped = true;
    for (int k = 0; k < p.clip_count; ++k) {
        const erd::DevClip& c = p.clips[k];
        double pl[8][3], E[3], cmin[3], cmax[3];
        for (int i = 0; i < 3; ++i) {
            E[i] = 1.0e-6 * (std::fabs((double)c.inv_tm[i][0]) * amax[0] + std::fabs((double)c.inv_tm[i][1]) * amax[1] +
                             std::fabs((double)c.inv_tm[i][2]) * amax[2] + std::fabs((double)c.inv_tm[i][3])) + 1.0e-30;
            cmin[i] = 1.0e300; cmax[i] = -1.0e300;
        }
        for (int q = 0; q < 8; ++q) {
            const double P[3] = {(q & 1) ? hi[0] : lo[0], (q & 2) ? hi[1] : lo[1], (q & 4) ? hi[2] : lo[2]};
            for (int i = 0; i < 3; ++i) {
                pl[q][i] = (double)c.inv_tm[i][0] * P[0] + (double)c.inv_tm[i][1] * P[1] + (double)c.inv_tm[i][2] * P[2] +
                           (double)c.inv_tm[i][3];
                cmin[i] = std::fmin(cmin[i], pl[q][i]);
                cmax[i] = std::fmax(cmax[i], pl[q][i]);
            }
        }
        bool all = true;
        for (int i = 0; i < 3; ++i) all = all && cmin[i] == cmin[i] && cmax[i] == cmax[i] && E[i] < 1.0e300;
        if (!all) { unclipped = false; continue; }
        const double h[3] = {0.5 * (double)c.size[0], 0.5 * (double)c.size[1], 0.5 * (double)c.size[2]};
        const double r = (double)c.outer_radius;
        /* the box inside the shape: the corners are inside with the margin */
        bool inside = false;
        if (c.type == ER_SHAPE_BOX) {
            inside = true;
            for (int i = 0; i < 3; ++i) inside = inside && cmax[i] < h[i] - E[i] && cmin[i] > -h[i] + E[i];
        } else if (c.type == ER_SHAPE_SPHERE) {
            const double lim = r - (std::sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]) + 1.0e-6 * std::fabs(r));
            inside = true;
            for (int q = 0; q < 8; ++q)
                inside = inside && std::sqrt(pl[q][0] * pl[q][0] + pl[q][1] * pl[q][1] + pl[q][2] * pl[q][2]) < lim;
        } else if (c.type == ER_SHAPE_CYLINDER) {
            const double lim = r - (std::sqrt(E[0] * E[0] + E[2] * E[2]) + 1.0e-6 * std::fabs(r));
            inside = cmax[1] < h[1] - E[1] && cmin[1] > -h[1] + E[1];
            for (int q = 0; q < 8; ++q) inside = inside && std::sqrt(pl[q][0] * pl[q][0] + pl[q][2] * pl[q][2]) < lim;
        } else {
            inside = cmin[2] > E[2];
        }
        /* the box separated from the shape */
        bool outside = false;
        if (c.type == ER_SHAPE_BOX) {
            for (int i = 0; i < 3; ++i) outside = outside || cmin[i] > h[i] + E[i] || cmax[i] < -h[i] - E[i];
        } else if (c.type == ER_SHAPE_SPHERE || c.type == ER_SHAPE_CYLINDER) {
            /* distance from the axis (cylinder: in x-z) or the centre to the corners' bounding box */
            double d2 = 0.0;
            for (int i = 0; i < 3; ++i) {
                if (c.type == ER_SHAPE_CYLINDER && i == 1) continue;
                const double g = std::fmax(0.0, std::fmax(cmin[i], -cmax[i]));
                d2 += g * g;
            }
            const double e = c.type == ER_SHAPE_SPHERE ? std::sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2])
                                                       : std::sqrt(E[0] * E[0] + E[2] * E[2]);
            outside = std::sqrt(d2) > r + e + 1.0e-6 * std::fabs(r);
            if (c.type == ER_SHAPE_CYLINDER) outside = outside || cmin[1] > h[1] + E[1] || cmax[1] < -h[1] - E[1];
        } else {
            outside = cmax[2] < -E[2];
        }
        if (c.invert ? outside : inside) return ER_MC_CLIPPED;
        if (!(c.invert ? inside : outside)) unclipped = false;
    }
    return unclipped ? ER_MC_UNCLIPPED : 0;
}

/* the clipping classes of every macrocell (mc_clip_class), on up to 16 host threads: gone bit m
 * set when macrocell m is wholly clipped away, free bit m when it is wholly unclipped */
void clip_classes(const DevParams& P, int nmc, double edge_ext, std::vector<uint32_t>& gone, std::vector<uint32_t>& free)
{
    const int words = (nmc + 31) / 32;
    gone.assign((size_t)words, 0u);
    free.assign((size_t)words, 0u);
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t]() {
            for (int w = t; w < words; w += nt) {
                uint32_t g = 0, f = 0;
                for (int b = 0; b < 32; ++b) {
                    const int m = w * 32 + b;
                    if (m >= nmc) break;
                    const int c = mc_clip_class(P, m, edge_ext);
                    if (c & ER_MC_CLIPPED) g |= 1u << b;
                    if (c & ER_MC_UNCLIPPED) f |= 1u << b;
                }
                gone[w] = g;
                free[w] = f;
            }
        });
    for (auto& x : th) x.join();
}

/* The empty-space radius of every group of macrocells from the groups' occupancy bits (dims:
 * the group grid), 4 bits each, eight to a word: 0 for an occupied group, else the smallest
 * r >= 1 such that the groups within Chebyshev distance r of it include an occupied one, at
 * most 15. Every group within distance r - 1 is then empty, so a march in that group may skip
 * the whole cube at once. The set of groups within distance r of an occupied group grows by
 * one group along each axis per round (a cube dilates as three axis segments). */
void group_radii(const std::vector<uint32_t>& occ, const int dims[3], std::vector<uint32_t>& out)
{
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    const size_t ng = (size_t)nx * ny * nz, sy = (size_t)nx, sz = (size_t)nx * ny;
    std::vector<uint8_t> cur(ng), tmp(ng), rad(ng);
    for (size_t g = 0; g < ng; ++g) {
        cur[g] = (uint8_t)((occ[g >> 5] >> (g & 31)) & 1u);
        rad[g] = cur[g] ? 0 : 15;
    }
    for (int r = 1; r < 15; ++r) {
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const size_t i = z * sz + y * sy + x;
                    tmp[i] = cur[i] | (x > 0 ? cur[i - 1] : 0) | (x < nx - 1 ? cur[i + 1] : 0);
                }
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const size_t i = z * sz + y * sy + x;
                    cur[i] = tmp[i] | (y > 0 ? tmp[i - sy] : 0) | (y < ny - 1 ? tmp[i + sy] : 0);
                }
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    const size_t i = z * sz + y * sy + x;
                    tmp[i] = cur[i] | (z > 0 ? cur[i - sz] : 0) | (z < nz - 1 ? cur[i + sz] : 0);
                }
        cur.swap(tmp);
        for (size_t g = 0; g < ng; ++g)
            if (rad[g] == 15 && cur[g]) rad[g] = (uint8_t)r;
    }
    out.assign((ng + 7) / 8, 0u);
    for (size_t g = 0; g < ng; ++g) out[g >> 3] |= (uint32_t)rad[g] << ((g & 7) * 4);
}

/* Rebuild the tracer's DevParams (and the LUT / macrocell maxima / gm_max that depend on it). */
int build_params(Tracer& t)
{
    if (t.built_version == g_scene_version) return ER_OK;
    int st;
    if ((st = ensure_neglog_device())) return st;
    const er_tracer_desc& d = t.desc;
    Volume* vol = g_volumes.find(d.volume_id);
    if (!vol) return set_error(ER_WARNING, "volume with ID:%d does not exist", d.volume_id);
    DevParams& P = t.host_params;
    memset(&P, 0, sizeof(P));
    camera_update(P, d.camera);
    const er_render_settings& s = d.settings;
    P.step_primary = s.step_factor_primary;
    P.step_shadow = s.step_factor_shadow;
    P.max_shadow = s.max_shadow_distance;
    P.density_scale = s.density_scale;
    P.grad_threshold = s.gradient_threshold;
    P.grad_factor = s.gradient_factor;
    P.shadows = s.shadows;
    P.estimator = s.estimator;
    P.shading_type = s.shading_type;
    P.opacity_modulated = s.opacity_modulated;
    P.gradient_mode = s.gradient_computation;
    P.sampler = s.sampler;

    /* Volume::operator=(ErVolume) (Source/volume.h:100-129) */
    const er_volume_desc& vd = vol->desc;
    float scale = 0.0f;
    if (vd.normalize_size) {
        const float ps[3] = {(float)vd.resolution[0] * vd.spacing[0], (float)vd.resolution[1] * vd.spacing[1],
                             (float)vd.resolution[2] * vd.spacing[2]};
        scale = 1.0f / std::max(ps[0], std::max(ps[1], ps[2]));
    }
    for (int i = 0; i < 3; ++i) {
        P.res[i] = vd.resolution[i];
        P.resf[i] = (float)vd.resolution[i];
        P.spacing[i] = vd.spacing[i] * scale;
        P.size[i] = (float)vd.resolution[i] * P.spacing[i];
        P.invsize[i] = 1.0f / P.size[i];
        P.minp[i] = P.size[i] * -0.5f;
        P.maxp[i] = P.size[i] * 0.5f;
    }
    P.minstep = std::min(P.spacing[0], std::min(P.spacing[1], P.spacing[2]));
    P.cells = vol->cells.ptr;
    P.cstride_y = vd.resolution[0] + 1;
    P.cstride_z = (vd.resolution[0] + 1) * (vd.resolution[1] + 1);
    P.neglog = g_neglog_dev;
    const int mc_shift = (t.mc_shift > 0 && brick_fits(vd, vol->vmax, t.mc_shift)) ? t.mc_shift : vol->mc_shift;
    MacroGrid* grid = volume_grid(*vol, mc_shift, st);
    if (!grid) return st;
    for (int i = 0; i < 3; ++i) P.mc_dim[i] = grid->mc[i];
    P.tile_shift = t.tile_shift;

    /* The tracer's lights, compacted into device slots in the order of its LightIDs, and the
     * textures they use, compacted likewise. The device indexes them by slot, so any number of
     * registered objects (er_set_list_capacity) and any IDs work; a tracer uses at most
     * ER_LIST_CAPACITY distinct lights and textures. */
    P.light_count = std::min(d.light_count, ER_MAX_INDICES);
    for (int i = 0; i < ER_MAX_INDICES; ++i) P.light_ids[i] = -1;
    std::map<int, int> light_slot, tex_slot;
    /* texture ID -> slot; unknown IDs evaluate black (textures.h:64-65) */
    auto texture_slot = [&](int id, int& slot) -> int {
        slot = -1;
        if (!g_textures.find(id)) return ER_OK;
        auto tt = tex_slot.find(id);
        if (tt == tex_slot.end()) {
            const int ts = (int)tex_slot.size();
            if (ts >= ER_LIST_CAPACITY) return set_error(ER_ERROR, "a tracer uses at most %d textures", ER_LIST_CAPACITY);
            tt = tex_slot.emplace(id, ts).first;
        }
        slot = tt->second;
        return ER_OK;
    };
    for (int i = 0; i < P.light_count; ++i) {
        LightObj* lo = g_lights.find(d.light_ids[i]);
        if (!lo) continue;
        auto it = light_slot.find(d.light_ids[i]);
        if (it != light_slot.end()) { P.light_ids[i] = it->second; continue; }
        const int slot = (int)light_slot.size();
        if (slot >= ER_LIST_CAPACITY) return set_error(ER_ERROR, "a tracer uses at most %d lights", ER_LIST_CAPACITY);
        light_slot[d.light_ids[i]] = slot;
        P.light_ids[i] = slot;
        const er_light_desc& L = lo->desc;
        erd::DevLight& D = P.lights[slot];
        mat34(D.tm, L.shape.tm);
        mat34(D.inv_tm, L.shape.inv_tm);
        D.type = L.shape.type; D.one_sided = L.shape.one_sided; D.visible = L.visible; D.unit = L.unit;
        for (int k = 0; k < 3; ++k) D.size[k] = L.shape.size[k];
        D.area = shape_area(L.shape);
        D.inner_radius = L.shape.inner_radius; D.outer_radius = L.shape.outer_radius;
        D.multiplier = L.multiplier; D.valid = 1;
        if ((st = texture_slot(L.texture_id, D.texture_id))) return st;
    }

    /* Surface objects (objects.h:46-82). Compat: the first ObjectIDs.Count live objects in ID
     * order, which is the device list the reference indexes with the loop counter (objects.h:52,77)
     * (clamped to the live count, where the reference would read past its list); full: the
     * objects the tracer's ObjectIDs name. Their textures join the texture slots. */
    std::vector<const er_object_desc*> objs;
    if (s.estimator == ER_ESTIMATOR_FULL) {
        for (int k = 0; k < d.object_count && k < ER_MAX_INDICES; ++k)
            if (ObjectObj* o = g_objects.find(d.object_ids[k])) objs.push_back(&o->desc);
    } else {
        for (auto& kv : g_objects.map) {
            if ((int)objs.size() >= d.object_count) break;
            objs.push_back(&kv.second->desc);
        }
    }
    if ((int)objs.size() > ER_LIST_CAPACITY) return set_error(ER_ERROR, "a tracer uses at most %d objects", ER_LIST_CAPACITY);
    P.object_count = (int)objs.size();
    for (int k = 0; k < P.object_count; ++k) {
        const er_object_desc& O = *objs[k];
        erd::DevObject& D = P.objects[k];
        erd::DevLight& S = D.shape;
        mat34(S.tm, O.shape.tm);
        mat34(S.inv_tm, O.shape.inv_tm);
        S.type = O.shape.type; S.one_sided = O.shape.one_sided; S.visible = 1; S.unit = 0;
        for (int i = 0; i < 3; ++i) S.size[i] = O.shape.size[i];
        S.area = shape_area(O.shape);
        S.inner_radius = O.shape.inner_radius; S.outer_radius = O.shape.outer_radius;
        S.multiplier = 0.0f; S.valid = 1; S.texture_id = -1;
        if ((st = texture_slot(O.diffuse_texture_id, D.diffuse_tex))) return st;
        if ((st = texture_slot(O.specular_texture_id, D.specular_tex))) return st;
        if ((st = texture_slot(O.glossiness_texture_id, D.gloss_tex))) return st;
        D.valid = 1;
    }
    int ngrad = 0;
    std::map<int, int> bitmap_slot;
    for (auto& kv : tex_slot) {
        const er_texture_desc& T = g_textures.find(kv.first)->desc;
        erd::DevTexture& D = P.textures[kv.second];
        D.type = T.type; D.proc_type = T.procedural.type; D.bitmap_id = -1; D.valid = 1;
        D.output_level = T.output_level;
        for (int i = 0; i < 2; ++i) { D.offset[i] = T.offset[i]; D.repeat[i] = T.repeat[i]; D.flip[i] = T.flip[i]; }
        for (int i = 0; i < 3; ++i) {
            D.uniform[i] = T.procedural.uniform_color[i];
            D.checker1[i] = T.procedural.checker_color1[i];
            D.checker2[i] = T.procedural.checker_color2[i];
        }
        D.gradient_index = -1;
        if (T.type == ER_TEXTURE_PROCEDURAL && T.procedural.type == ER_PROCEDURAL_GRADIENT) {
            if (ngrad >= 4) return set_error(ER_ERROR, "at most 4 gradient textures are supported");
            P.tex_gradients[ngrad] = T.procedural.gradient;
            D.gradient_index = ngrad++;
        }
        /* bitmaps (textures.h:99-105): uploaded once per set of pixels, compacted into slots */
        BitmapObj* bo = T.type == ER_TEXTURE_BITMAP ? g_bitmaps.find(T.bitmap_id) : nullptr;
        if (bo && !bo->pixels.empty()) {
            auto bt = bitmap_slot.find(T.bitmap_id);
            if (bt == bitmap_slot.end()) {
                const int bs = (int)bitmap_slot.size();
                if (bs >= ER_LIST_CAPACITY) return set_error(ER_ERROR, "a tracer uses at most %d bitmaps", ER_LIST_CAPACITY);
                const size_t texels = bo->pixels.size() / 4;
                if ((size_t)std::max(0, bo->desc.resolution[0]) * std::max(0, bo->desc.resolution[1]) != texels)
                    return set_error(ER_ERROR, "bitmap %d: resolution does not match its pixels", T.bitmap_id);
                if (bo->dev_stale || bo->dev.n != texels) {
                    if ((st = bo->dev.alloc(texels))) return st;
                    HIP_TRY(hipMemcpy(bo->dev.ptr, bo->pixels.data(), texels * sizeof(uchar4), hipMemcpyHostToDevice));
                    bo->dev_stale = false;
                }
                erd::DevBitmap& B = P.bitmaps[bs];
                B.pixels = bo->dev.ptr;
                B.w = bo->desc.resolution[0];
                B.h = bo->desc.resolution[1];
                bt = bitmap_slot.emplace(T.bitmap_id, bs).first;
            }
            D.bitmap_id = bt->second;
        }
    }
    P.clip_count = 0;
    if (s.estimator == ER_ESTIMATOR_FULL) {
        for (int k = 0; k < d.clipping_count && k < ER_MAX_INDICES; ++k) {
            ClipObj* c = g_clips.find(d.clipping_ids[k]);
            if (!c || P.clip_count >= ER_LIST_CAPACITY) continue;
            erd::DevClip& D = P.clips[P.clip_count++];
            mat34(D.inv_tm, c->desc.shape.inv_tm);
            D.type = c->desc.shape.type; D.invert = c->desc.invert;
            for (int i = 0; i < 3; ++i) D.size[i] = c->desc.shape.size[i];
            D.outer_radius = c->desc.shape.outer_radius;
        }
    }

    P.opacity = d.opacity; P.glossiness = d.glossiness;
    P.diffuse = d.diffuse; P.specular = d.specular; P.emission = d.emission;

    /* node-table transfer function */
    P.tf_lut = nullptr;
    P.tf_lut_scale = t.tf_scale;
    if (!t.tf_lut.empty()) {
        if ((st = t.tf_lut_dev.alloc(t.tf_lut.size()))) return st;
        HIP_TRY(hipMemcpy(t.tf_lut_dev.ptr, t.tf_lut.data(), t.tf_lut.size() * sizeof(float), hipMemcpyHostToDevice));
        P.tf_lut = t.tf_lut_dev.ptr;
    }

    /* exact opacity LUT over the intensities the volume can produce */
    const int lut_n = (int)vol->vmax + 1;
    std::vector<float> lut(lut_n);
    for (int i = 0; i < lut_n; ++i) {
        const float o = t.tf_lut.empty() ? plf_eval_host(d.opacity, (float)i)
                                         : erd::tf_lut_eval(t.tf_lut.data(), t.tf_scale, (float)i, 0);
        lut[i] = s.density_scale * o;
    }
    if ((st = t.lut.alloc(lut_n))) return st;
    HIP_TRY(hipMemcpy(t.lut.ptr, lut.data(), lut_n * sizeof(float), hipMemcpyHostToDevice));
    P.sigma_lut = t.lut.ptr;
    P.lut_max = lut_n - 1;
    const int nmc = grid->mc[0] * grid->mc[1] * grid->mc[2];
    if ((st = t.mc_sigma.alloc(nmc))) return st;
    HIP_TRY(erd_launch_macro_sigma(grid->mm.ptr, t.lut.ptr, P.lut_max, t.mc_sigma.ptr, nmc, t.stream));
    P.mc_sigma = t.mc_sigma.ptr;
    P.mc_shift = mc_shift;
    /* Woodcock majorants on their own fixed grid of 2^ER_WC_SHIFT-cell macrocells */
    P.wc_sigma = nullptr;
    if (s.estimator == ER_ESTIMATOR_FULL && s.sampler == ER_SAMPLER_WOODCOCK) {
        MacroGrid* wg = volume_grid(*vol, ER_WC_SHIFT, st);
        if (!wg) return st;
        const int nw = wg->mc[0] * wg->mc[1] * wg->mc[2];
        if ((st = t.wc_sigma.alloc(nw))) return st;
        HIP_TRY(erd_launch_macro_sigma(wg->mm.ptr, t.lut.ptr, P.lut_max, t.wc_sigma.ptr, nw, t.stream));
        P.wc_sigma = t.wc_sigma.ptr;
        for (int i = 0; i < 3; ++i) P.wc_dim[i] = wg->mc[i];
    }
    /* occupancy bits: one per macrocell; then the radii of the groups of 2^ER_GROUP_SHIFT
     * macrocells a side (4 bits each) and the groups' clip-free bits, the part the march kernels
     * stage in LDS; then the groups' occupancy bits the radii are computed from */
    const int mc_words = (nmc + 31) / 32;
    for (int i = 0; i < 3; ++i) P.mc2_dim[i] = (grid->mc[i] + (1 << ER_GROUP_SHIFT) - 1) >> ER_GROUP_SHIFT;
    const int ng = P.mc2_dim[0] * P.mc2_dim[1] * P.mc2_dim[2];
    const int g_words = (ng + 31) / 32, r_words = (ng + 7) / 8;
    P.mc2_base = mc_words;
    P.mc3_base = mc_words + r_words;
    P.mc_words = mc_words + r_words + g_words;
    const int occ_base = P.mc_words;
    if ((st = t.mc_bits.alloc(P.mc_words + g_words))) return st;
    P.mc_bits = t.mc_bits.ptr;
    t.lut_lds = (size_t)lut_n * sizeof(float) <= 32768u;
    t.lds_bytes = (size_t)P.mc_words * 4 + (t.lut_lds ? (size_t)lut_n * sizeof(float) : 0);

    if ((st = t.params.alloc(1))) return st;
    P.gm_inv = 0.0f;
    t.gm_max = 0.0f;
    HIP_TRY(hipMemcpyAsync(t.params.ptr, &P, sizeof(P), hipMemcpyHostToDevice, t.stream));
    /* macrocells wholly clipped away count as empty (the bits kernel reads the clipping objects
     * from the parameters just uploaded) */
    const uint32_t* clipped = nullptr;
    const uint32_t* unclipped = nullptr;
    if (P.clip_count > 0) {
        std::vector<uint32_t> gone, free;
        clip_classes(P, nmc, ray_origin_bound(P, objs) * 1.0e-5, gone, free);
        gone.insert(gone.end(), free.begin(), free.end());
        if ((st = t.clip_bits.alloc(gone.size()))) return st;
        HIP_TRY(hipMemcpyAsync(t.clip_bits.ptr, gone.data(), gone.size() * sizeof(uint32_t), hipMemcpyHostToDevice, t.stream));
        HIP_TRY(hipStreamSynchronize(t.stream));
        clipped = t.clip_bits.ptr;
        unclipped = t.clip_bits.ptr + mc_words;
    }
    HIP_TRY(erd_launch_macro_bits(t.mc_sigma.ptr, clipped, t.mc_bits.ptr, nmc, mc_words, t.stream));
    HIP_TRY(erd_launch_group_bits(t.mc_bits.ptr, unclipped, grid->mc[0], grid->mc[1], grid->mc[2], P.mc2_dim[0],
                                  P.mc2_dim[1], P.mc2_dim[2], occ_base, P.mc3_base, g_words, t.stream));
    {
        std::vector<uint32_t> occ((size_t)g_words), radii;
        HIP_TRY(hipMemcpyAsync(occ.data(), t.mc_bits.ptr + occ_base, occ.size() * sizeof(uint32_t), hipMemcpyDeviceToHost,
                               t.stream));
        HIP_TRY(hipStreamSynchronize(t.stream));
        group_radii(occ, P.mc2_dim, radii);
        HIP_TRY(hipMemcpyAsync(t.mc_bits.ptr + P.mc2_base, radii.data(), radii.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice, t.stream));
        HIP_TRY(hipStreamSynchronize(t.stream));
    }
    if (needs_gm(d)) {
        if ((st = t.gm_tmp.alloc(1))) return st;
        HIP_TRY(hipMemsetAsync(t.gm_tmp.ptr, 0, sizeof(unsigned int), t.stream));
        HIP_TRY(erd_launch_gm_max(t.params.ptr, t.gm_tmp.ptr, t.stream));
        unsigned int bits = 0;
        HIP_TRY(hipMemcpyAsync(&bits, t.gm_tmp.ptr, sizeof(bits), hipMemcpyDeviceToHost, t.stream));
        HIP_TRY(hipStreamSynchronize(t.stream));
        memcpy(&t.gm_max, &bits, 4);
        P.gm_inv = t.gm_max > 0.0f ? 1.0f / t.gm_max : 0.0f;
        HIP_TRY(hipMemcpyAsync(t.params.ptr, &P, sizeof(P), hipMemcpyHostToDevice, t.stream));
    }
    /* bilateral tables (Source/filter.h:34-40) */
    if (s.estimator == ER_ESTIMATOR_FULL && s.bilateral) {
        const int R = s.bilateral_radius;
        if (R < 0 || R > 32) return set_error(ER_ERROR, "bilateral radius must be in [0, 32]");
        std::vector<char> bp(erd_bilateral_params_size(), 0);
        int* radius = (int*)bp.data();
        float* kd = (float*)(bp.data() + sizeof(int));
        float* gs = kd + 65;
        *radius = R;
        const float ss = s.bilateral_sigma_spatial, sr = s.bilateral_sigma_range;
        for (int k = -R; k <= R; ++k) kd[k + R] = erm::expf_(-((float)(k * k)) / (2.0f * ss * ss));
        for (int q = 0; q < 256; ++q) gs[q] = erm::expf_(-((float)(q * q)) / (2.0f * sr * sr));
        if ((st = t.bilateral.alloc(bp.size()))) return st;
        HIP_TRY(hipMemcpyAsync(t.bilateral.ptr, bp.data(), bp.size(), hipMemcpyHostToDevice, t.stream));
    }
    HIP_TRY(hipStreamSynchronize(t.stream));
    t.built_version = g_scene_version;
    return ER_OK;
}

bool use_bilateral(const Tracer& t)
{
    return t.desc.settings.estimator == ER_ESTIMATOR_FULL && t.desc.settings.bilateral;
}

int primary_blocks(const Tracer& t) { return t.blocks_per_cu > 0 ? g_cus * t.blocks_per_cu : g_primary_blocks; }
int shadow_blocks(const Tracer& t) { return std::max(1, g_cus * t.shadow_waves_per_cu / 4); }

/* One pass on the tracer's two streams (no host synchronisation). Timing events, when given:
 * e[0] first stream before the primary kernel, e[1] after it, e[2] second stream after the
 * combine kernel, e[3] after post. */
int enqueue_pass(Tracer& t, hipEvent_t* e)
{
    if (t.W == 0 || t.H == 0) return ER_OK;
    const int b = 1 - t.cur;
    PassSet& S = t.set[b];
    if (!e) e = S.marks;
    const int full = t.desc.settings.estimator == ER_ESTIMATOR_FULL;
    const int shade_blocks = std::max(1, std::min((t.W * t.H + 255) / 256, 2048));
    /* the set was last used two passes ago: its post must have consumed the frame and queues */
    if (S.pending) HIP_TRY(hipStreamWaitEvent(t.stream, S.posted, 0));
    if (e) HIP_TRY(hipEventRecord(e[0], t.stream));
    const int split = 1;   /* every sampler defers its shadow marches (Woodcock's run on their own stream) */
    /* Woodcock tracking is per lane: its marches take one lane per ray, primary and shadow */
    const bool woodcock = full && t.desc.settings.sampler == ER_SAMPLER_WOODCOCK;
    const size_t lds_bits = (size_t)t.host_params.mc_words * 4;
    const int ntiles = ((t.W + (1 << t.tile_shift) - 1) >> t.tile_shift) * ((t.H + (64 >> t.tile_shift) - 1) / (64 >> t.tile_shift));
    if (t.tile_cost.n < (size_t)ntiles) {
        /* every pass writes the cost of each of its tiles before the order kernel reads them */
        int st;
        if ((st = sync_tracer(t)) || (st = t.tile_cost.alloc(ntiles)) || (st = t.tile_order.alloc(ntiles + 1))) return st;
        t.order_tiles = 0;
    }
    const bool order_valid = t.order_shift == t.tile_shift && t.order_tiles == ntiles;
    /* tile costs change slowly from pass to pass: the order is rebuilt every fourth pass (and
     * whenever it does not fit the tile shape) */
#ifndef ER_REORDER_MASK
#define ER_REORDER_MASK 3
#endif
    const bool reorder = !order_valid || (t.desc.no_iterations & ER_REORDER_MASK) == 0;
    HIP_TRY(erd_launch_trace(t.params.ptr, S.frame.ptr, t.seeds1.ptr, t.seeds2.ptr, S.queue.ptr, S.qcount.ptr,
                             S.events.ptr, S.jobs.ptr, S.vis.ptr, t.counters.ptr, t.W, t.H, full, t.counting, t.lut_lds,
                             split, t.lds_bytes, lds_bits, shade_blocks, primary_blocks(t), shadow_blocks(t), t.tile_shift, t.stream,
                             t.stream2, S.shaded, e ? e[1] : nullptr, order_valid ? t.tile_order.ptr : nullptr,
                             t.tile_cost.ptr, reorder ? t.tile_order.ptr : nullptr, t.heavy_tiles,
                             woodcock ? 1 : 64 / t.rays_per_wave, woodcock ? 1 : 64 / t.shadow_rays_per_wave,
                             64 / t.heavy_rays_per_wave));
    t.order_shift = t.tile_shift;
    t.order_tiles = ntiles;
    if (e) HIP_TRY(hipEventRecord(e[2], t.stream2));
    HIP_TRY(erd_launch_post(t.params.ptr, S.frame.ptr, t.running.ptr, t.display.ptr, t.desc.no_iterations,
                            S.qcount.ptr, t.W, t.H, t.stream2));
    if (use_bilateral(t))
        HIP_TRY(erd_launch_bilateral(t.bilateral.ptr, t.display.ptr, t.disp_tmp.ptr, t.disp_filt.ptr, t.W, t.H, t.stream2));
    HIP_TRY(hipEventRecord(S.posted, t.stream2));
    S.pending = true;
    if (e) HIP_TRY(hipEventRecord(e[3], t.stream2));
    t.cur = b;
    t.desc.no_iterations++;   /* core.cu:135 */
    return ER_OK;
}

Tracer* get_tracer(int id, int& st)
{
    Tracer* t = g_tracers.find(id);
    st = t ? ER_OK : set_error(ER_WARNING, "resource item with ID:%d does not exist", id);
    return t;
}

int copy_out(void* dst, const void* src, size_t bytes, size_t size, hipStream_t s)
{
    if (!dst) return set_error(ER_ERROR, "output pointer is NULL");
    if (size < bytes) return set_error(ER_ERROR, "output buffer too small (%zu < %zu bytes)", size, bytes);
    if (bytes == 0) return ER_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ER_OK;
}

template<class R, class D, class F>
int bind_generic(R& reg, D* d, int bind, F make)
{
    if (!d) return set_error(ER_ERROR, "descriptor is NULL");
    if (!bind) {
        auto it = reg.map.find(d->id);
        if (d->id < 0 || it == reg.map.end()) return ER_OK;   /* list.cuh:89-93 */
        delete it->second;
        reg.map.erase(it);
        g_scene_version++;
        return ER_OK;
    }
    if (reg.map.size() + 1 >= (size_t)g_list_capacity) return ER_OK;   /* list.cuh:63-67 */
    auto* obj = reg.find(d->id);
    const bool fresh = obj == nullptr;
    if (fresh) {
        const int id = reg.counter;
        if (id >= (1 << 30)) return set_error(ER_ERROR, "ID space exhausted");
        d->id = id;
    }
    int st = make(obj, *d, fresh);
    if (st) {
        if (fresh) d->id = -1;   /* nothing was bound */
        return st;
    }
    if (fresh) { reg.map[d->id] = obj; reg.counter++; }
    g_scene_version++;
    return ER_OK;
}

/* ---- autotuner ------------------------------------------------------------------------ */

/* What a tuned configuration depends on: the tracer's descriptor (film, camera, transfer
 * functions, settings, lights, objects), its node table, and the volume's contents. */
std::string tune_key(const Tracer& t)
{
    er_tracer_desc d = t.desc;
    d.id = -1;
    d.no_iterations = 0;
    std::string k((const char*)&d, sizeof(d));
    const Volume* v = g_volumes.find(t.desc.volume_id);
    const uint64_t up = v ? v->upload : 0;
    k.append((const char*)&up, sizeof(up));
    k.append((const char*)t.tf_lut.data(), t.tf_lut.size() * sizeof(float));
    k.append((const char*)&t.tf_scale, sizeof(float));
    return k;
}

std::map<std::string, er_tune> g_tune_cache;

void fill_tune(const Tracer& t, er_tune& r)
{
    const Volume* v = g_volumes.find(t.desc.volume_id);
    const int shift = (t.mc_shift > 0 && v && brick_fits(v->desc, v->vmax, t.mc_shift)) ? t.mc_shift : (v ? v->mc_shift : 3);
    r.brick = 1 << shift;
    r.tile_w = 1 << t.tile_shift;
    r.tile_h = 64 >> t.tile_shift;
    r.blocks_per_cu = t.blocks_per_cu > 0 ? t.blocks_per_cu : g_primary_blocks / std::max(1, g_cus);
    r.shadow_waves_per_cu = t.shadow_waves_per_cu;
    r.heavy_tiles = t.heavy_tiles;
    r.rays_per_wave = t.rays_per_wave;
    r.shadow_rays_per_wave = t.shadow_rays_per_wave;
    r.heavy_rays_per_wave = t.heavy_rays_per_wave;
}

int apply_tune(Tracer& t, const er_tune& r)
{
    int bs = 0, ts = 0;
    while (bs < 8 && (1 << bs) != r.brick) ++bs;
    while (ts < 7 && (1 << ts) != r.tile_w) ++ts;
    if ((1 << bs) != r.brick || bs < 1 || bs > 6) return set_error(ER_ERROR, "brick must be 2, 4, ..., 64 cells");
    if ((1 << ts) != r.tile_w || ts < 2 || ts > 5 || r.tile_w * r.tile_h != 64)
        return set_error(ER_ERROR, "tile must be 4x16, 8x8, 16x4 or 32x2 pixels");
    if (r.blocks_per_cu < 1 || r.blocks_per_cu > 16) return set_error(ER_ERROR, "blocks per CU must be in [1, 16]");
    if (r.shadow_waves_per_cu < 1 || r.shadow_waves_per_cu > 64)
        return set_error(ER_ERROR, "shadow waves per CU must be in [1, 64]");
    if (r.heavy_tiles < 0 || r.heavy_tiles > 4096) return set_error(ER_ERROR, "heavy tiles must be in [0, 4096]");
    if (r.rays_per_wave != 64 && r.rays_per_wave != 16 && r.rays_per_wave != 4 && r.rays_per_wave != 1)
        return set_error(ER_ERROR, "rays per wave must be 64, 16, 4 or 1");
    if (r.shadow_rays_per_wave != 64 && r.shadow_rays_per_wave != 16 && r.shadow_rays_per_wave != 4)
        return set_error(ER_ERROR, "shadow rays per wave must be 64, 16 or 4");
    if (r.heavy_rays_per_wave != 16 && r.heavy_rays_per_wave != 4 && r.heavy_rays_per_wave != 1)
        return set_error(ER_ERROR, "heavy-tile rays per wave must be 16, 4 or 1");
    const Volume* v = g_volumes.find(t.desc.volume_id);
    if (v && !brick_fits(v->desc, v->vmax, bs))
        return set_error(ER_ERROR, "a %d-cell brick's occupancy bits and opacity table exceed %zu bytes of LDS", r.brick,
                         ER_MAX_TABLES_BYTES);
    t.mc_shift = bs;
    t.tile_shift = ts;
    t.blocks_per_cu = r.blocks_per_cu;
    t.shadow_waves_per_cu = r.shadow_waves_per_cu;
    t.heavy_tiles = r.heavy_tiles;
    t.rays_per_wave = r.rays_per_wave;
    t.shadow_rays_per_wave = r.shadow_rays_per_wave;
    t.heavy_rays_per_wave = r.heavy_rays_per_wave;
    t.built_version = 0;
    return ER_OK;
}

/* The per-tracer state a timing trial overwrites, saved and put back so that tuning leaves
 * a render where it was. */
struct Snapshot {
    DevBuf<uint32_t> s1, s2;
    DevBuf<float4> run;
    DevBuf<uchar4> disp, filt;
    DevBuf<unsigned long long> cnt;
    int no_iter = 0;
    ~Snapshot() { s1.release(); s2.release(); run.release(); disp.release(); filt.release(); cnt.release(); }
    int save(Tracer& t)
    {
        int st;
        const size_t n = (size_t)t.W * t.H;
        if ((st = sync_tracer(t))) return st;
        if ((st = s1.alloc(n)) || (st = s2.alloc(n)) || (st = run.alloc(n)) || (st = disp.alloc(n)) ||
            (st = filt.alloc(n)) || (st = cnt.alloc(12)))
            return st;
        HIP_TRY(hipMemcpy(s1.ptr, t.seeds1.ptr, n * 4, hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(s2.ptr, t.seeds2.ptr, n * 4, hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(run.ptr, t.running.ptr, n * sizeof(float4), hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(disp.ptr, t.display.ptr, n * sizeof(uchar4), hipMemcpyDeviceToDevice));
        /* the bilateral filter's output is what er_get_estimate returns when it is on */
        HIP_TRY(hipMemcpy(filt.ptr, t.disp_filt.ptr, n * sizeof(uchar4), hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(cnt.ptr, t.counters.ptr, 12 * sizeof(unsigned long long), hipMemcpyDeviceToDevice));
        no_iter = t.desc.no_iterations;
        return ER_OK;
    }
    /* the seeds only: every trial traces the same paths */
    int rewind(Tracer& t)
    {
        const size_t n = (size_t)t.W * t.H;
        HIP_TRY(hipMemcpyAsync(t.seeds1.ptr, s1.ptr, n * 4, hipMemcpyDeviceToDevice, t.stream));
        HIP_TRY(hipMemcpyAsync(t.seeds2.ptr, s2.ptr, n * 4, hipMemcpyDeviceToDevice, t.stream));
        return ER_OK;
    }
    int restore(Tracer& t)
    {
        int st;
        const size_t n = (size_t)t.W * t.H;
        if ((st = sync_tracer(t))) return st;
        HIP_TRY(hipMemcpy(t.seeds1.ptr, s1.ptr, n * 4, hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(t.seeds2.ptr, s2.ptr, n * 4, hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(t.running.ptr, run.ptr, n * sizeof(float4), hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(t.display.ptr, disp.ptr, n * sizeof(uchar4), hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(t.disp_filt.ptr, filt.ptr, n * sizeof(uchar4), hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(t.counters.ptr, cnt.ptr, 12 * sizeof(unsigned long long), hipMemcpyDeviceToDevice));
        t.desc.no_iterations = no_iter;
        return ER_OK;
    }
};

/* wall time of `passes` passes from the saved seeds (first stream start to last post end),
 * after one warm-up pass */
int time_passes(Tracer& t, Snapshot& snap, int passes, float& ms)
{
    int st;
    if ((st = build_params(t))) return st;
    if ((st = snap.rewind(t)) || (st = enqueue_pass(t, nullptr)) || (st = sync_tracer(t)) || (st = snap.rewind(t)))
        return st;
    if ((st = timing_events(t, (size_t)passes * 4))) return st;
    hipEvent_t* ev = t.timing_events.data();
    for (int k = 0; k < passes && !st; ++k) st = enqueue_pass(t, &ev[4 * k]);
    const int sst = sync_tracer(t);
    if (!st) st = sst;
    if (!st) {
        hipError_t e = hipEventElapsedTime(&ms, ev[0], ev[4 * (passes - 1) + 3]);
        if (e != hipSuccess) st = set_error(ER_ERROR, "%s (hipEventElapsedTime)", hipGetErrorString(e));
    }
    return st;
}

/* The candidates of one search axis around `cur`: 0 bricks that fit every volume, 1 tile
 * shapes, 2 workgroups per CU, 3 shadow waves per CU, 4 heavy tiles (with 64 rays per wave
 * only), 5 rays per wave, 6 shadow rays per wave, 7 heavy-tile rays per wave (with heavy tiles
 * only). */
std::vector<er_tune> tune_candidates(int axis, const er_tune& cur, const std::vector<const Volume*>& vols)
{
    std::vector<er_tune> c;
    if (axis == 0) {
        for (int b = 1; b <= 5; ++b) {
            bool fits = true;
            for (const Volume* v : vols) fits = fits && v && brick_fits(v->desc, v->vmax, b);
            if (fits && (1 << b) != cur.brick) { er_tune x = cur; x.brick = 1 << b; c.push_back(x); }
        }
    } else if (axis == 1) {
        for (int ts = 2; ts <= 5; ++ts)
            if ((1 << ts) != cur.tile_w) { er_tune x = cur; x.tile_w = 1 << ts; x.tile_h = 64 >> ts; c.push_back(x); }
    } else if (axis == 2) {
        for (int b = 3; b <= 9; ++b)
            if (b != cur.blocks_per_cu) { er_tune x = cur; x.blocks_per_cu = b; c.push_back(x); }
    } else if (axis == 3) {
        for (int w = 2; w <= 8; w *= 2)
            if (w != cur.shadow_waves_per_cu) { er_tune x = cur; x.shadow_waves_per_cu = w; c.push_back(x); }
    } else if (axis == 4) {
        /* heavy tiles together with their lanes per ray: a count of heavy tiles pays off only
         * with the right lane count, which the later axis alone would not find from the default */
        if (cur.rays_per_wave == 64)
            for (int h : {0, 256, 768})
                for (int r : {16, 4, 1}) {
                    if (h == 0 && r != cur.heavy_rays_per_wave) continue;
                    if (h == cur.heavy_tiles && (h == 0 || r == cur.heavy_rays_per_wave)) continue;
                    er_tune x = cur;
                    x.heavy_tiles = h;
                    x.heavy_rays_per_wave = r;
                    c.push_back(x);
                }
    } else if (axis == 7) {
        if (cur.rays_per_wave == 64 && cur.heavy_tiles > 0)
            for (int r : {16, 4, 1})
                if (r != cur.heavy_rays_per_wave) { er_tune x = cur; x.heavy_rays_per_wave = r; c.push_back(x); }
    } else if (axis == 5) {
        for (int r : {16, 4, 1})
            if (r != cur.rays_per_wave) { er_tune x = cur; x.rays_per_wave = r; c.push_back(x); }
    } else {
        for (int r : {64, 16, 4})
            if (r != cur.shadow_rays_per_wave) { er_tune x = cur; x.shadow_rays_per_wave = r; c.push_back(x); }
    }
    return c;
}

/* a candidate replaces the incumbent only when it is clearly faster: timing noise must not
 * walk the search away from a good configuration */
const float ER_TUNE_MARGIN = 0.97f;
const float ER_TUNE_FINE_MARGIN = 0.985f;   /* the closer look's margin (median of six trials) */
const float ER_TUNE_TRIAL_MS = 30.0f;   /* target length of one timing trial */

} /* namespace */

/* ============================== C ABI ================================================== */
extern "C" {

ER_API int er_abi_version(void) { return 2; }

ER_API int er_abi_sizeof(int* sizes, int n)
{
    const int s[13] = {(int)sizeof(er_tracer_desc), (int)sizeof(er_volume_desc), (int)sizeof(er_light_desc),
                       (int)sizeof(er_texture_desc), (int)sizeof(er_clipping_desc), (int)sizeof(er_object_desc),
                       (int)sizeof(er_bitmap_desc), (int)sizeof(er_plf), (int)sizeof(er_shape),
                       (int)sizeof(er_timing), (int)sizeof(er_counters), (int)sizeof(er_batch_item),
                       (int)sizeof(er_tune)};
    int k = 0;
    for (; sizes && k < n && k < 13; ++k) sizes[k] = s[k];
    return k;
}

ER_API void er_plf_init(er_plf* p)
{
    memset(p, 0, sizeof(*p));
    p->node_range[0] = FLT_MAX;
    p->node_range[1] = FLT_MIN;
}

ER_API void er_plf_add_node(er_plf* p, float position, float value)
{
    if (p->count + 1 >= ER_MAX_NO_TF_NODES) return;
    p->position[p->count] = position;
    p->value[p->count] = value;
    if (position < p->node_range[0]) p->node_range[0] = position;
    if (position > p->node_range[1]) p->node_range[1] = position;
    p->count++;
}

static void shape_init(er_shape* s)
{
    memset(s, 0, sizeof(*s));
    for (int i = 0; i < 4; ++i) { s->tm[i][i] = 1.0f; s->inv_tm[i][i] = 1.0f; }
    s->type = ER_SHAPE_PLANE;
}

ER_API void er_tracer_desc_init(er_tracer_desc* d)
{
    memset(d, 0, sizeof(*d));
    d->id = -1;
    d->enabled = 1;
    er_plf_init(&d->opacity);
    er_plf_init(&d->glossiness);
    for (int c = 0; c < 3; ++c) { er_plf_init(&d->diffuse.plf[c]); er_plf_init(&d->specular.plf[c]); er_plf_init(&d->emission.plf[c]); }
    er_render_settings& s = d->settings;
    s.step_factor_primary = 0.1f; s.step_factor_shadow = 0.1f; s.shadows = 1; s.max_shadow_distance = 1.0f;
    s.shading_type = ER_SHADING_BRDF_ONLY; s.density_scale = 100.0f; s.opacity_modulated = 1;
    s.gradient_computation = ER_GRADIENT_CENTRAL; s.gradient_threshold = 0.5f; s.gradient_factor = 0.5f;
    s.estimator = ER_ESTIMATOR_COMPAT; s.sampler = ER_SAMPLER_RAYMARCH;
    s.bilateral = 0; s.bilateral_radius = 3; s.bilateral_sigma_spatial = 2.0f; s.bilateral_sigma_range = 25.0f;
    for (int i = 0; i < ER_MAX_INDICES; ++i) { d->light_ids[i] = -1; d->object_ids[i] = -1; d->clipping_ids[i] = -1; }
}

ER_API void er_volume_desc_init(er_volume_desc* d)
{
    memset(d, 0, sizeof(*d));
    d->id = -1; d->enabled = 1;
    d->spacing[0] = d->spacing[1] = d->spacing[2] = 1.0f;
}

ER_API void er_light_desc_init(er_light_desc* d)
{
    memset(d, 0, sizeof(*d));
    d->id = -1; d->enabled = 1; d->texture_id = -1; d->unit = ER_UNIT_POWER;
    shape_init(&d->shape);
}

ER_API void er_texture_desc_init(er_texture_desc* d)
{
    memset(d, 0, sizeof(*d));
    d->id = -1; d->enabled = 1; d->type = ER_TEXTURE_PROCEDURAL; d->output_level = 1.0f; d->bitmap_id = -1;
    for (int c = 0; c < 3; ++c) er_plf_init(&d->procedural.gradient.plf[c]);
}

ER_API void er_clipping_desc_init(er_clipping_desc* d)
{
    memset(d, 0, sizeof(*d));
    d->id = -1; d->enabled = 1;
    shape_init(&d->shape);
}

ER_API void er_object_desc_init(er_object_desc* d)
{
    memset(d, 0, sizeof(*d));
    d->id = -1; d->enabled = 1; d->diffuse_texture_id = d->specular_texture_id = d->glossiness_texture_id = -1;
    d->ior = 5.0f;
    shape_init(&d->shape);
}

ER_API void er_bitmap_desc_init(er_bitmap_desc* d)
{
    memset(d, 0, sizeof(*d));
    d->id = -1; d->enabled = 1;
}

ER_API int er_last_error(char* buf, size_t size)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    if (buf && size) {
        strncpy(buf, g_last_error.c_str(), size - 1);
        buf[size - 1] = 0;
    }
    return (int)g_last_error.size();
}

ER_API int er_init(void)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st = ensure_device();
    if (st) return st;
    return ensure_neglog_device();
}

ER_API int er_shutdown(void)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    g_tracers.clear(); g_volumes.clear(); g_lights.clear(); g_textures.clear(); g_clips.clear();
    g_objects.clear(); g_bitmaps.clear();
    if (g_neglog_dev) { (void)hipFree(g_neglog_dev); g_neglog_dev = nullptr; }
    clear_batch_graphs();
    if (g_batch_stream) { (void)hipStreamDestroy(g_batch_stream); g_batch_stream = nullptr; }
    if (g_batch_stream2) { (void)hipStreamDestroy(g_batch_stream2); g_batch_stream2 = nullptr; }
    g_scene_version++;
    return ER_OK;
}

ER_API int er_synchronize(void)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    if (!g_device_ready) return ER_OK;
    HIP_TRY(hipDeviceSynchronize());
    return ER_OK;
}

ER_API int er_bind_tracer(er_tracer_desc* d, int bind)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    if (bind) { int st = ensure_device(); if (st) return st; }
    return bind_generic(g_tracers, d, bind, [](Tracer*& t, er_tracer_desc& desc, bool fresh) -> int {
        if (fresh) {
            t = new Tracer();
            hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipStreamCreateWithFlags(&t->stream2, hipStreamNonBlocking);
            if (e != hipSuccess) { delete t; t = nullptr; return set_error(ER_ERROR, "%s (hipStreamCreate)", hipGetErrorString(e)); }
            int st = t->counters.alloc(12);
            if (!st && hipMemset(t->counters.ptr, 0, 12 * sizeof(unsigned long long)) != hipSuccess)
                st = set_error(ER_ERROR, "hipMemset failed");
            if (st) { delete t; t = nullptr; return st; }
        }
        t->desc = desc;
        int st = tracer_resize(*t, desc.camera.film_size[0], desc.camera.film_size[1], desc.id);
        if (st && fresh) { delete t; t = nullptr; }
        return st;
    });
}

ER_API int er_bind_volume(er_volume_desc* d, int bind)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    if (bind) { int st = ensure_device(); if (st) return st; }
    return bind_generic(g_volumes, d, bind, [](Volume*& v, er_volume_desc& desc, bool fresh) -> int {
        if (fresh) v = new Volume();
        const bool upload = fresh || desc.voxels_dirty;
        /* without new voxels the device buffers keep their size: the volume keeps its own
         * resolution and takes the rest (spacing, normalisation) from the descriptor, as
         * Buffer3D::operator= skips a source that is not Dirty (buffer3d.h:52-65) */
        int res[3] = {v->desc.resolution[0], v->desc.resolution[1], v->desc.resolution[2]};
        v->desc = desc;
        if (!upload)
            for (int i = 0; i < 3; ++i) v->desc.resolution[i] = res[i];
        if (upload) {
            int st = upload_volume(*v);
            if (st) { if (fresh) { delete v; v = nullptr; } return st; }
            desc.voxels_dirty = 0;   /* Buffer3D dirty-gated copy (buffer3d.h:52-65) */
        }
        v->desc.voxels = nullptr;
        return ER_OK;
    });
}

ER_API int er_bind_light(er_light_desc* d, int bind)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    return bind_generic(g_lights, d, bind, [](LightObj*& o, er_light_desc& desc, bool fresh) -> int {
        if (fresh) o = new LightObj();
        desc.shape.area = shape_area(desc.shape);   /* Light::operator= -> Shape::Update (light.h:39-45) */
        o->desc = desc;
        return ER_OK;
    });
}

ER_API int er_bind_object(er_object_desc* d, int bind)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    return bind_generic(g_objects, d, bind, [](ObjectObj*& o, er_object_desc& desc, bool fresh) -> int {
        if (fresh) o = new ObjectObj();
        desc.shape.area = shape_area(desc.shape);
        o->desc = desc;
        return ER_OK;
    });
}

ER_API int er_bind_clipping_object(er_clipping_desc* d, int bind)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    return bind_generic(g_clips, d, bind, [](ClipObj*& o, er_clipping_desc& desc, bool fresh) -> int {
        if (fresh) o = new ClipObj();
        o->desc = desc;
        return ER_OK;
    });
}

ER_API int er_bind_texture(er_texture_desc* d, int bind)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    return bind_generic(g_textures, d, bind, [](TextureObj*& o, er_texture_desc& desc, bool fresh) -> int {
        if (fresh) o = new TextureObj();
        o->desc = desc;
        return ER_OK;
    });
}

ER_API int er_bind_bitmap(er_bitmap_desc* d, int bind)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    return bind_generic(g_bitmaps, d, bind, [](BitmapObj*& o, er_bitmap_desc& desc, bool fresh) -> int {
        if (fresh) o = new BitmapObj();
        const size_t n = (size_t)std::max(0, desc.resolution[0]) * std::max(0, desc.resolution[1]) * 4;
        if ((fresh || desc.pixels_dirty) && desc.pixels && n) {
            o->pixels.assign(desc.pixels, desc.pixels + n);
            o->dev_stale = true;
            desc.pixels_dirty = 0;
        }
        o->desc = desc;
        o->desc.pixels = nullptr;
        return ER_OK;
    });
}

ER_API int er_render_estimate(int tracer_id)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    if ((st = build_params(*t))) return st;
    if ((st = enqueue_pass(*t, nullptr))) return st;
    if ((st = sync_tracer(*t))) return st;
    return ER_OK;
}

ER_API int er_render_passes(int tracer_id, int passes, er_timing* timing)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    if (passes < 0) return set_error(ER_ERROR, "negative pass count");
    if ((st = build_params(*t))) return st;
    if (timing && (st = timing_events(*t, (size_t)passes * 4))) return st;
    hipEvent_t* ev = t->timing_events.data();
    for (int k = 0; k < passes && !st; ++k) st = enqueue_pass(*t, timing ? &ev[4 * k] : nullptr);
    hipError_t se = hipStreamSynchronize(t->stream);
    hipError_t se2 = hipStreamSynchronize(t->stream2);
    if (se == hipSuccess) se = se2;
    if (timing && !st && se == hipSuccess) {
        float total = 0, tr = 0, po = 0, pr = 0, ms = 0;
        for (int k = 0; k < passes; ++k) {
            if (hipEventElapsedTime(&ms, ev[4 * k], ev[4 * k + 1]) == hipSuccess) pr += ms;
            if (hipEventElapsedTime(&ms, ev[4 * k], ev[4 * k + 2]) == hipSuccess) tr += ms;
            if (hipEventElapsedTime(&ms, ev[4 * k + 2], ev[4 * k + 3]) == hipSuccess) po += ms;
        }
        if (passes > 0 && hipEventElapsedTime(&ms, ev[0], ev[4 * (passes - 1) + 3]) == hipSuccess) total = ms;
        timing->total_ms = total; timing->trace_ms = tr; timing->post_ms = po; timing->passes = passes;
        timing->primary_ms = pr;
    }
    if (st) return st;
    if (se != hipSuccess) return set_error(ER_ERROR, "%s (hipStreamSynchronize)", hipGetErrorString(se));
    return ER_OK;
}

/* Coordinate search over brick size, tile shape and persistent workgroups per CU. Each
 * candidate renders the same passes from the same seeds (median of three trials of about
 * 30 ms, or `passes` passes when given); a candidate must be 3% faster to be taken. The
 * tracer's render state is restored afterwards. Results are cached by scene. */
ER_API int er_autotune(int tracer_id, int passes, int flags, er_tune* out)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* tp = get_tracer(tracer_id, st);
    if (!tp) return st;
    Tracer& t = *tp;
    if ((st = build_params(t))) return st;
    const std::string key = tune_key(t);
    auto cached = g_tune_cache.find(key);
    if (cached != g_tune_cache.end() && !(flags & ER_TUNE_FORCE)) {
        er_tune r = cached->second;
        if ((st = apply_tune(t, r))) return st;
        if ((st = build_params(t))) return st;
        r.trials = 0;
        if (out) *out = r;
        return ER_OK;
    }
    if (t.W == 0 || t.H == 0) return set_error(ER_ERROR, "tracer %d has an empty film", tracer_id);
    Snapshot snap;
    if ((st = snap.save(t))) return st;
    const int mc0 = t.mc_shift, ts0 = t.tile_shift, bpc0 = t.blocks_per_cu, sw0 = t.shadow_waves_per_cu,
              ht0 = t.heavy_tiles, rw0 = t.rays_per_wave, srw0 = t.shadow_rays_per_wave, hrw0 = t.heavy_rays_per_wave;
    er_tune cur{};
    fill_tune(t, cur);
    int trials = 0;
    int reps = 3;
    /* a candidate's time is the median of its trials: some configurations (many shadow waves
     * per CU) vary by 20% from trial to trial, and the best of their trials ranked them ahead
     * of steadier ones that render long runs faster */
    auto measure = [&](const er_tune& c, float& med_ms) -> int {
        int s;
        if ((s = apply_tune(t, c))) return s;
        float ms[8];
        for (int rep = 0; rep < reps; ++rep)
            if ((s = time_passes(t, snap, passes, ms[rep]))) return s;
        std::sort(ms, ms + reps);
        med_ms = (reps & 1) ? ms[reps / 2] : 0.5f * (ms[reps / 2 - 1] + ms[reps / 2]);
        ++trials;
        return ER_OK;
    };
    float cur_ms = 0.0f;
    if (passes <= 0) {
        /* size the trials from one probe pass of the starting configuration */
        float probe = 0.0f;
        passes = 2;
        if (!(st = time_passes(t, snap, passes, probe)))
            passes = std::max(2, std::min(64, (int)std::ceil(ER_TUNE_TRIAL_MS / std::max(probe / 2.0f, 1e-3f))));
    }
    if (!st) st = measure(cur, cur_ms);
    const float default_ms = cur_ms;
    std::vector<const Volume*> vols{g_volumes.find(t.desc.volume_id)};
    /* rays per wave right after the brick: the later axes are tuned around them. A second sweep,
     * from a fresh measurement of the incumbent, catches axes that interact (the primary
     * workgroups and the shadow waves share the CUs); it is skipped when the first moved
     * nothing. */
    for (int sweep = 0; sweep < 2 && !st; ++sweep) {
        bool moved = false;
        if (sweep > 0 && (st = measure(cur, cur_ms))) break;
        for (int axis : {0, 5, 6, 1, 2, 3, 4, 7}) {
            if (st) break;
            for (const er_tune& c : tune_candidates(axis, cur, vols)) {
                float ms = 0.0f;
                if ((st = measure(c, ms))) break;
                if (ms < ER_TUNE_MARGIN * cur_ms) { cur = c; cur_ms = ms; moved = true; }
            }
        }
        if (!moved) break;
    }
    /* a closer look at the axes whose values lie within a few percent of each other (tile
     * shape, primary workgroups, shadow waves per CU): median of six trials instead of three,
     * and a candidate need only be 1.5% faster */
    if (!st) {
        reps = 6;
        st = measure(cur, cur_ms);
        for (int axis : {1, 2, 3}) {
            if (st) break;
            for (const er_tune& c : tune_candidates(axis, cur, vols)) {
                float ms = 0.0f;
                if ((st = measure(c, ms))) break;
                if (ms < ER_TUNE_FINE_MARGIN * cur_ms) { cur = c; cur_ms = ms; }
            }
        }
    }
    if (!st) st = apply_tune(t, cur);
    if (st) {
        t.mc_shift = mc0; t.tile_shift = ts0; t.blocks_per_cu = bpc0; t.shadow_waves_per_cu = sw0; t.heavy_tiles = ht0;
        t.rays_per_wave = rw0; t.shadow_rays_per_wave = srw0; t.heavy_rays_per_wave = hrw0;
        t.built_version = 0;
    }
    const int rst = snap.restore(t);
    if (st) return st;
    if (rst) return rst;
    if ((st = build_params(t))) return st;
    cur.ms_per_pass = cur_ms / (float)passes;
    cur.default_ms_per_pass = default_ms / (float)passes;
    cur.trials = trials;
    g_tune_cache[key] = cur;
    if (out) *out = cur;
    return ER_OK;
}

/* One configuration for a whole batch: the same coordinate search, each candidate applied to
 * every job (a brick is a candidate only if it fits every job's volume) and timed as `rounds`
 * batch rounds from the jobs' saved seeds (best of three). Every job's render state is
 * restored; the jobs keep the chosen configuration. ms_per_pass reports one round. */
ER_API int er_autotune_batch(const er_batch_item* items, int n, int rounds, er_tune* out)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    if (n <= 0 || !items) return set_error(ER_ERROR, "empty batch");
    if (rounds <= 0) rounds = 4;
    std::vector<Tracer*> ts;
    std::vector<const Volume*> vols;
    for (int k = 0; k < n; ++k) {
        Tracer* t = get_tracer(items[k].tracer_id, st);
        if (!t) return st;
        if ((st = build_params(*t))) return st;
        ts.push_back(t);
        vols.push_back(g_volumes.find(t->desc.volume_id));
    }
    std::vector<std::unique_ptr<Snapshot>> snaps;
    std::vector<er_tune> orig(n);
    for (int k = 0; k < n; ++k) {
        snaps.emplace_back(new Snapshot());
        if ((st = snaps[k]->save(*ts[k]))) return st;
        fill_tune(*ts[k], orig[k]);
    }
    std::vector<er_batch_item> trial(items, items + n);
    for (auto& it : trial) it.passes = rounds;
    int trials = 0;
    auto apply_all = [&](const er_tune& c) -> int {
        int s;
        for (int k = 0; k < n; ++k) {
            er_tune x = c;
            if (vols[k] && !brick_fits(vols[k]->desc, vols[k]->vmax, (int)std::log2((double)c.brick))) x.brick = orig[k].brick;
            if ((s = apply_tune(*ts[k], x))) return s;
        }
        return ER_OK;
    };
    auto measure = [&](const er_tune& c, float& best_ms) -> int {
        int s;
        if ((s = apply_all(c))) return s;
        best_ms = FLT_MAX;
        for (int rep = 0; rep < 4; ++rep) {   /* the first is a warm-up */
            for (int k = 0; k < n; ++k)
                if ((s = snaps[k]->rewind(*ts[k])) || (s = sync_tracer(*ts[k]))) return s;
            er_timing tm{};
            if ((s = er_render_batch(trial.data(), n, &tm))) return s;
            if (rep > 0) best_ms = std::min(best_ms, tm.total_ms);
        }
        ++trials;
        return ER_OK;
    };
    er_tune cur = orig[0];
    cur.brick = orig[0].brick;
    float cur_ms = 0.0f;
    st = measure(cur, cur_ms);
    const float default_ms = cur_ms;
    for (int axis : {0, 5, 6, 1, 2}) {
        if (st) break;
        for (const er_tune& c : tune_candidates(axis, cur, vols)) {
            float ms = 0.0f;
            if ((st = measure(c, ms))) break;
            if (ms < ER_TUNE_MARGIN * cur_ms) { cur = c; cur_ms = ms; }
        }
    }
    if (!st) st = apply_all(cur);
    if (st) for (int k = 0; k < n; ++k) (void)apply_tune(*ts[k], orig[k]);
    for (int k = 0; k < n; ++k) {
        const int rst = snaps[k]->restore(*ts[k]);
        if (!st) st = rst;
    }
    if (st) return st;
    cur.ms_per_pass = cur_ms / (float)rounds;
    cur.default_ms_per_pass = default_ms / (float)rounds;
    cur.trials = trials;
    if (out) *out = cur;
    return ER_OK;
}

ER_API int er_set_tune(int tracer_id, const er_tune* tune)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    if (!tune) return set_error(ER_ERROR, "tune is NULL");
    if ((st = sync_tracer(*t))) return st;
    return apply_tune(*t, *tune);
}

ER_API int er_get_tune(int tracer_id, er_tune* tune)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    if (!tune) return set_error(ER_ERROR, "tune is NULL");
    er_tune r{};
    fill_tune(*t, r);
    *tune = r;
    return ER_OK;
}


ER_API int er_set_list_capacity(int capacity)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    if (capacity < 2 || capacity > (1 << 20)) return set_error(ER_ERROR, "capacity must be in [2, 2^20]");
    g_list_capacity = capacity;
    return ER_OK;
}

/* One graph of batch rounds over jobs sorted by pass count (at most ER_BATCH_CONST_JOBS). */
static int render_batch_chunk(std::vector<std::pair<int, Tracer*>>& jobs, int full, er_timing* timing)
{
    int st;
    const int nj = (int)jobs.size();
    /* two job tables, one per buffer set: rounds alternate between the sets so that a round's
     * shadow marches, combine and post (second stream) overlap the next round's primary and
     * shading (first stream) */
    std::vector<erd::BatchJob> table[2];
    int tile_base = 0, post_base = 0;
    for (int b = 0; b < 2; ++b) table[b].resize(nj);
    for (int k = 0; k < nj; ++k) {
        Tracer& t = *jobs[k].second;
        for (int b = 0; b < 2; ++b) {
            PassSet& S = t.set[b];
            erd::BatchJob& B = table[b][k];
            B.params = t.params.ptr;
            B.frame = S.frame.ptr; B.running = t.running.ptr; B.display = t.display.ptr;
            B.seeds1 = t.seeds1.ptr; B.seeds2 = t.seeds2.ptr;
            B.queue = (erd::ScatterRec*)S.queue.ptr; B.events = (erd::EventRec*)S.events.ptr;
            B.jobs = (erd::ShadowJob*)S.jobs.ptr; B.vis = S.vis.ptr;
            B.counts = S.qcount.ptr;
            B.tiles_x = (t.W + (1 << t.tile_shift) - 1) >> t.tile_shift;
            B.ntiles = B.tiles_x * ((t.H + (64 >> t.tile_shift) - 1) / (64 >> t.tile_shift));
            B.tile_base = tile_base;
            B.post_tiles_x = (t.W + 15) / 16;
            B.post_ntiles = B.post_tiles_x * ((t.H + 15) / 16);
            B.post_base = post_base;
            B.passes = jobs[k].first;
            B.no_iter0 = t.desc.no_iterations;
        }
        tile_base += table[0][k].ntiles;
        post_base += table[0][k].post_ntiles;
    }
    if (!g_batch_stream) HIP_TRY(hipStreamCreateWithFlags(&g_batch_stream, hipStreamNonBlocking));
    if (!g_batch_stream2) HIP_TRY(hipStreamCreateWithFlags(&g_batch_stream2, hipStreamNonBlocking));
    hipStream_t bs = g_batch_stream, bs2 = g_batch_stream2;
    const int shade_blocks = 2048;
    int batch_blocks = 0;   /* persistent primary workgroups: the largest any job asks for */
    /* LDS per primary / shadow workgroup: the largest job's macrocell bits plus opacity table,
     * as a single pass stages them; when some job's opacity table is too large for LDS, every
     * job reads its tables from global memory (lds_words = 0) */
    int lds_words = 0;
    bool all_lds = true;
    for (auto& j : jobs) {
        const Tracer& t = *j.second;
        batch_blocks = std::max(batch_blocks, primary_blocks(t));
        lds_words = std::max(lds_words, (int)(t.lds_bytes / 4));
        all_lds = all_lds && t.lut_lds;
    }
    if (!all_lds) lds_words = 0;
    /* lanes per primary and shadow march, for the whole batch: the jobs' common rays per wave
     * (one lane per ray when they differ, or when a job uses Woodcock tracking) */
    int lpr = 64 / jobs[0].second->rays_per_wave, slpr = 64 / jobs[0].second->shadow_rays_per_wave;
    for (auto& j : jobs) {
        const Tracer& t = *j.second;
        const bool wc = full && t.desc.settings.sampler == ER_SAMPLER_WOODCOCK;
        if (wc || 64 / t.rays_per_wave != lpr) lpr = 1;
        if (wc || 64 / t.shadow_rays_per_wave != slpr) slpr = 1;
    }
    const int rounds = jobs[0].first;
    /* the graph's key: the job tables without their first pass numbers, and the launch shape */
    std::string key;
    {
        const int hdr[7] = {nj, full, batch_blocks, lds_words, shade_blocks, lpr, slpr};
        key.append((const char*)hdr, sizeof(hdr));
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < nj; ++k) {
                erd::BatchJob B = table[b][k];
                B.no_iter0 = 0;
                key.append((const char*)&B, sizeof(B));
            }
    }
    auto cached = g_batch_graphs.find(key);
    const bool hit = cached != g_batch_graphs.end();
    if (!hit && g_batch_graphs.size() >= ER_BATCH_GRAPH_CACHE) clear_batch_graphs();
    BatchGraph& G = hit ? cached->second : g_batch_graphs[key];
    DevBuf<erd::BatchJob>* dJ = G.dJ;
    for (int b = 0; b < 2; ++b) {
        if ((st = dJ[b].alloc(nj))) { G.release(); g_batch_graphs.erase(key); return st; }
        HIP_TRY(hipMemcpy(dJ[b].ptr, table[b].data(), nj * sizeof(erd::BatchJob), hipMemcpyHostToDevice));
        for (int k = 0; k < nj; ++k) HIP_TRY(hipMemset(table[b][k].counts, 0, 4 * sizeof(uint32_t)));
    }
    /* the jobs' parameter blocks, in table order, for the batch kernels */
    {
        std::vector<erd::DevParams> hp(nj);
        for (int k = 0; k < nj; ++k) hp[k] = jobs[k].second->host_params;
        HIP_TRY(erd_batch_params_upload(hp.data(), nj));
    }
    /* all rounds as one graph (no memset nodes): per round, primary + shading on the first
     * stream, then shadow + combine + post + queue reset on the second; a set is reused two
     * rounds later, after its post */
    std::vector<hipEvent_t> evs;
    auto new_event = [&](hipEvent_t& e) -> hipError_t {
        hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
        if (r == hipSuccess) evs.push_back(e);
        return r;
    };
    hipGraph_t graph = G.graph;
    hipGraphExec_t exec = G.exec;
    hipError_t ec = hit ? hipSuccess : hipStreamBeginCapture(bs, hipStreamCaptureModeThreadLocal);
    if (!hit) {
        hipEvent_t fork = nullptr, posted[2] = {nullptr, nullptr};
        if (ec == hipSuccess) ec = new_event(fork);
        if (ec == hipSuccess) ec = hipEventRecord(fork, bs);
        if (ec == hipSuccess) ec = hipStreamWaitEvent(bs2, fork, 0);
        for (int r = 0; r < rounds && ec == hipSuccess; ++r) {
            const int b = r & 1;
            /* the jobs taking part in round r: those with more than r passes, a prefix of the
             * table (ordered by pass count, largest first), and their post tiles */
            int na = 0, active_tiles = 0;
            while (na < nj && jobs[na].first > r) active_tiles += table[0][na++].post_ntiles;
            if (posted[b]) ec = hipStreamWaitEvent(bs, posted[b], 0);
            hipEvent_t shaded = nullptr;
            if (ec == hipSuccess) ec = new_event(shaded);
            if (ec == hipSuccess)
                ec = erd_launch_batch_trace(dJ[b].ptr, na, full, batch_blocks, lds_words, shade_blocks, bs, bs2, shaded, lpr,
                                            slpr);
            if (ec == hipSuccess) ec = erd_launch_batch_post(dJ[b].ptr, na, (unsigned int)r, active_tiles, bs2);
            if (ec == hipSuccess) ec = erd_launch_round_end(dJ[b].ptr, nj, bs2);
            if (ec == hipSuccess) ec = new_event(posted[b]);
            if (ec == hipSuccess) ec = hipEventRecord(posted[b], bs2);
        }
        hipEvent_t join = nullptr;
        if (ec == hipSuccess) ec = new_event(join);
        if (ec == hipSuccess) ec = hipEventRecord(join, bs2);
        if (ec == hipSuccess) ec = hipStreamWaitEvent(bs, join, 0);
        hipError_t ee = hipStreamEndCapture(bs, &graph);
        if (ec == hipSuccess) ec = ee;
        if (ec == hipSuccess) ec = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        for (auto e : evs) (void)hipEventDestroy(e);
        G.graph = graph;
        G.exec = exec;
        if (ec != hipSuccess) {
            G.release();
            g_batch_graphs.erase(key);
            return set_error(ER_ERROR, "batch graph: %s", hipGetErrorString(ec));
        }
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timing) {
        HIP_TRY(hipEventCreate(&t0));
        HIP_TRY(hipEventCreate(&t1));
        HIP_TRY(hipEventRecord(t0, bs));
    }
    hipError_t el = hipGraphLaunch(exec, bs);
    if (timing && el == hipSuccess) el = hipEventRecord(t1, bs);
    const hipError_t es = hipStreamSynchronize(bs);
    float ms = 0.0f;
    if (timing && el == hipSuccess && es == hipSuccess) (void)hipEventElapsedTime(&ms, t0, t1);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (el != hipSuccess) return set_error(ER_ERROR, "%s (hipGraphLaunch)", hipGetErrorString(el));
    if (es != hipSuccess) return set_error(ER_ERROR, "%s (hipStreamSynchronize)", hipGetErrorString(es));
    for (auto& j : jobs) {
        Tracer& t = *j.second;
        t.desc.no_iterations += j.first;
        t.cur = (j.first - 1) & 1;   /* the set of the tracer's last round */
        t.set[0].pending = t.set[1].pending = false;
        /* the bilateral display filter only shapes the displayed image: once, after the last pass */
        if (use_bilateral(t))
            HIP_TRY(erd_launch_bilateral(t.bilateral.ptr, t.display.ptr, t.disp_tmp.ptr, t.disp_filt.ptr, t.W, t.H, t.stream));
    }
    for (auto& j : jobs) if ((st = sync_tracer(*j.second))) return st;
    if (timing) {
        timing->total_ms = ms; timing->trace_ms = ms; timing->post_ms = 0.0f; timing->passes = rounds;
        timing->primary_ms = 0.0f;
    }
    return ER_OK;
}

ER_API int er_render_batch(const er_batch_item* items, int n, er_timing* timing)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    if (n <= 0 || !items) return set_error(ER_ERROR, "empty batch");
    std::vector<std::pair<int, Tracer*>> jobs;   /* (passes, tracer) */
    int full = -1;
    for (int k = 0; k < n; ++k) {
        Tracer* t = get_tracer(items[k].tracer_id, st);
        if (!t) return st;
        if (items[k].passes < 0) return set_error(ER_ERROR, "negative pass count");
        for (auto& j : jobs)
            if (j.second == t) return set_error(ER_ERROR, "tracer %d appears twice in the batch", items[k].tracer_id);
        const er_render_settings& s = t->desc.settings;
        const int f = s.estimator == ER_ESTIMATOR_FULL;
        if (full >= 0 && f != full) return set_error(ER_ERROR, "all tracers of a batch must use the same estimator mode");
        full = f;
        if ((st = build_params(*t))) return st;
        if ((st = sync_tracer(*t))) return st;
        if (items[k].passes > 0 && t->W > 0 && t->H > 0) jobs.push_back({items[k].passes, t});
    }
    if (jobs.empty()) return ER_OK;
    std::stable_sort(jobs.begin(), jobs.end(), [](const std::pair<int, Tracer*>& a, const std::pair<int, Tracer*>& b) {
        return a.first > b.first;
    });
    /* one graph per ER_BATCH_CONST_JOBS jobs of similar pass counts */
    er_timing sum{};
    for (size_t off = 0; off < jobs.size(); off += ER_BATCH_CONST_JOBS) {
        std::vector<std::pair<int, Tracer*>> chunk(jobs.begin() + off,
                                                   jobs.begin() + std::min(jobs.size(), off + ER_BATCH_CONST_JOBS));
        er_timing tm{};
        if ((st = render_batch_chunk(chunk, full, timing ? &tm : nullptr))) return st;
        sum.total_ms += tm.total_ms;
        sum.trace_ms += tm.trace_ms;
        sum.passes = std::max(sum.passes, tm.passes);
    }
    if (timing) *timing = sum;
    return ER_OK;
}

ER_API int er_set_tf_lut(int tracer_id, const float* lut, float range)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    if (lut) {
        if (!(range > 0.0f) || !std::isfinite(range)) return set_error(ER_ERROR, "node-table range must be positive");
        t->tf_lut.assign(lut, lut + ER_TF_LUT_NODES * 4);
        t->tf_scale = (float)(ER_TF_LUT_NODES - 1) / range;
    } else {
        t->tf_lut.clear();
        t->tf_scale = 0.0f;
    }
    t->desc.no_iterations = 0;
    ++g_scene_version;
    return ER_OK;
}

/* checks shared by the adjoint entry points; builds the tracer's parameters */
static int adjoint_setup(Tracer& t, const uint32_t* s1, const uint32_t* s2, size_t count, int passes)
{
    int st;
    if (t.tf_lut.empty()) return set_error(ER_ERROR, "the adjoint needs a node-table transfer function (er_set_tf_lut)");
    const er_render_settings& s = t.desc.settings;
    if (s.estimator != ER_ESTIMATOR_FULL || s.shading_type != ER_SHADING_PHASE_ONLY || s.sampler != ER_SAMPLER_RAYMARCH)
        return set_error(ER_ERROR, "the adjoint supports the full estimator with phase-function shading and the ray-march sampler");
    if (passes < 1) return set_error(ER_ERROR, "the adjoint needs at least one pass");
    if (!s1 || !s2) return set_error(ER_ERROR, "seed pointer is NULL");
    if (count < (size_t)t.W * t.H) return set_error(ER_ERROR, "need %zu seeds, got %zu", (size_t)t.W * t.H, count);
    if ((st = build_params(t))) return st;
    if (!t.lut_lds) return set_error(ER_ERROR, "the adjoint needs the opacity table in LDS (volume maximum <= 8191)");
    if (t.host_params.object_count > 0) return set_error(ER_ERROR, "the adjoint does not support surface objects");
    return ER_OK;
}

ER_API int er_adjoint_lut(int tracer_id, const uint32_t* seeds1, const uint32_t* seeds2, size_t count, int passes,
                          const float* dL_drunning_xyz, double* dL_dlut)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    if ((st = adjoint_setup(*t, seeds1, seeds2, count, passes))) return st;
    if (!dL_drunning_xyz || !dL_dlut) return set_error(ER_ERROR, "adjoint pointer is NULL");
    const size_t n = (size_t)t->W * t->H;
    const int blocks = ER_ADJ_BLOCKS;
    const size_t per_block = erd_adj_partials_per_block();
    size_t rs = 0, re = 0, rj = 0;
    erd_pass_record_sizes(&rs, &re, &rj);
    /* two buffer sets: pass k + 1's primary and shading kernels (first stream) overlap pass k's
     * adjoint kernels (second stream), as the forward passes do */
    struct AdjSet {
        DevBuf<uint32_t> s1s, s2s, counts, lcount;
        DevBuf<float> tvis;
        DevBuf<float4> frame;
        DevBuf<char> queue, events, jobs;
        DevBuf<uint2> list;
        hipEvent_t shaded = nullptr, done = nullptr;
        bool pending = false;
        ~AdjSet()
        {
            if (shaded) (void)hipEventDestroy(shaded);
            if (done) (void)hipEventDestroy(done);
        }
    } S[2];
    DevBuf<uint32_t> d1, d2;
    DevBuf<float> dldr, partials;
    DevBuf<float4> adj;
    DevBuf<double> out;
    if ((st = d1.alloc(n)) || (st = d2.alloc(n)) || (st = dldr.alloc(n * 3)) || (st = adj.alloc(n)) ||
        (st = partials.alloc(3 * (size_t)blocks * per_block)) || (st = out.alloc(per_block)))
        return st;
    for (AdjSet& A : S) {
        if ((st = A.s1s.alloc(n)) || (st = A.s2s.alloc(n)) || (st = A.counts.alloc(4)) || (st = A.lcount.alloc(1)) ||
            (st = A.tvis.alloc(2 * n)) || (st = A.frame.alloc(n)) || (st = A.queue.alloc(n * rs)) ||
            (st = A.events.alloc(n * re)) || (st = A.jobs.alloc(2 * n * rj)) || (st = A.list.alloc(n)))
            return st;
        HIP_TRY(hipEventCreateWithFlags(&A.shaded, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&A.done, hipEventDisableTiming));
        HIP_TRY(hipMemset(A.counts.ptr, 0, 4 * sizeof(uint32_t)));
        HIP_TRY(hipMemset(A.lcount.ptr, 0, sizeof(uint32_t)));
    }
    hipStream_t s = t->stream, s2 = t->stream2;
    HIP_TRY(hipMemcpy(d1.ptr, seeds1, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d2.ptr, seeds2, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dldr.ptr, dL_drunning_xyz, n * 12, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(out.ptr, 0, per_block * sizeof(double)));
    const int shade_blocks = std::max(1, std::min((int)((n + 255) / 256), 2048));
    const size_t lds_bits = (size_t)t->host_params.mc_words * 4;
    float cur_w = -1.0f;
    /* the running estimate after N passes is the mean of the filtered frames of passes 2..N
     * (the first pass is discarded, Q1), or the first frame when N == 1 */
    for (int k = 1; k <= passes; ++k) {
        AdjSet& A = S[k & 1];
        const float w = passes == 1 ? 1.0f : (k >= 2 ? 1.0f / (float)(passes - 1) : 0.0f);
        if (A.pending) HIP_TRY(hipStreamWaitEvent(s, A.done, 0));
        if (w != 0.0f) {
            HIP_TRY(hipMemcpyAsync(A.s1s.ptr, d1.ptr, n * 4, hipMemcpyDeviceToDevice, s));
            HIP_TRY(hipMemcpyAsync(A.s2s.ptr, d2.ptr, n * 4, hipMemcpyDeviceToDevice, s));
        }
        /* the forward pass's primary and shading kernels: decisions, queues, final seeds */
        HIP_TRY(erd_launch_primary_shade(t->params.ptr, A.frame.ptr, d1.ptr, d2.ptr, A.queue.ptr, A.counts.ptr, A.events.ptr,
                                         A.jobs.ptr, t->W, t->H, t->lut_lds, t->lds_bytes, lds_bits, shade_blocks,
                                         primary_blocks(*t), t->tile_shift, s));
        HIP_TRY(hipEventRecord(A.shaded, s));
        HIP_TRY(hipStreamWaitEvent(s2, A.shaded, 0));
        if (w != 0.0f) {
            if (w != cur_w) {
                HIP_TRY(erd_launch_adj_filter_t(dldr.ptr, adj.ptr, t->W, t->H, w, s2));
                cur_w = w;
            }
            HIP_TRY(erd_launch_adj_wavefront(t->params.ptr, A.counts.ptr, A.jobs.ptr, A.events.ptr, A.queue.ptr, adj.ptr,
                                             A.tvis.ptr, A.frame.ptr, A.list.ptr, A.lcount.ptr, A.s1s.ptr, A.s2s.ptr,
                                             partials.ptr, blocks, (int)n, t->lds_bytes, s2));
            HIP_TRY(erd_launch_adj_reduce(partials.ptr, 3 * blocks, out.ptr, s2));
        }
        HIP_TRY(erd_launch_adj_reset(A.counts.ptr, A.lcount.ptr, s2));
        HIP_TRY(hipEventRecord(A.done, s2));
        A.pending = true;
    }
    HIP_TRY(hipMemcpyAsync(dL_dlut, out.ptr, per_block * sizeof(double), hipMemcpyDeviceToHost, s2));
    HIP_TRY(hipStreamSynchronize(s2));
    if ((st = sync_tracer(*t))) return st;
    return ER_OK;
}

/* er_adjoint_lut with every work buffer in the device scratch array: no hipMalloc or hipFree.
 * The same kernels run for one pass at a time on the tracer's first stream (er_adjoint_lut
 * overlaps consecutive passes on two buffer sets). */
ER_API int er_adjoint_lut_noalloc(int tracer_id, const uint32_t* seeds1, const uint32_t* seeds2, size_t count, int passes,
                                  const float* dL_drunning_xyz, double* dL_dlut)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    NoAllocScope no_alloc;
    const size_t n = (size_t)t->W * t->H;
    if (n > ER_NOALLOC_MAX_PIXELS)
        return set_error(ER_ERROR, "er_adjoint_lut_noalloc: a film of %zu pixels exceeds the scratch limit of %zu pixels "
                                   "(512 x 512); use er_adjoint_lut", n, (size_t)ER_NOALLOC_MAX_PIXELS);
    if ((st = adjoint_setup(*t, seeds1, seeds2, count, passes))) return st;
    if (!dL_drunning_xyz || !dL_dlut) return set_error(ER_ERROR, "adjoint pointer is NULL");
    const size_t per_block = erd_adj_partials_per_block();
    if (per_block != (size_t)ER_TF_LUT_NODES * 4) return set_error(ER_ERROR, "adjoint histogram size mismatch");
    unsigned char* base = nullptr;
    if ((st = scratch_for_current_device(base))) return st;
    const NoAllocLayout L = noalloc_layout(n);
    uint32_t* d1 = (uint32_t*)(base + L.d1);
    uint32_t* d2 = (uint32_t*)(base + L.d2);
    uint32_t* s1s = (uint32_t*)(base + L.s1s);
    uint32_t* s2s = (uint32_t*)(base + L.s2s);
    float* dldr = (float*)(base + L.dldr);
    float4* adj = (float4*)(base + L.adj);
    float* tvis = (float*)(base + L.tvis);
    float4* frame = (float4*)(base + L.frame);
    void* queue = base + L.queue;
    void* events = base + L.events;
    void* jobs = base + L.jobs;
    uint2* list = (uint2*)(base + L.list);
    unsigned int* counts = (unsigned int*)(base + L.counts);
    unsigned int* lcount = (unsigned int*)(base + L.lcount);
    float* partials = (float*)(base + L.partials);
    double* out = (double*)(base + L.out);
    const int blocks = ER_ADJ_BLOCKS;
    hipStream_t s = t->stream;
    if ((st = sync_tracer(*t))) return st;
    HIP_TRY(hipMemcpy(d1, seeds1, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d2, seeds2, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dldr, dL_drunning_xyz, n * 12, hipMemcpyHostToDevice));
    /* the array keeps the previous call's contents: the queue counters and the gradient sum are
     * cleared by kernels (no memset nodes, should a caller capture the sequence) */
    HIP_TRY(erd_launch_adj_reset(counts, lcount, s));
    HIP_TRY(erd_launch_zero_words((uint32_t*)out, per_block * sizeof(double) / 4, s));
    const int shade_blocks = std::max(1, std::min((int)((n + 255) / 256), 2048));
    const size_t lds_bits = (size_t)t->host_params.mc_words * 4;
    float cur_w = -1.0f;
    for (int k = 1; k <= passes; ++k) {
        const float w = passes == 1 ? 1.0f : (k >= 2 ? 1.0f / (float)(passes - 1) : 0.0f);
        if (w != 0.0f) {
            HIP_TRY(hipMemcpyAsync(s1s, d1, n * 4, hipMemcpyDeviceToDevice, s));
            HIP_TRY(hipMemcpyAsync(s2s, d2, n * 4, hipMemcpyDeviceToDevice, s));
        }
        HIP_TRY(erd_launch_primary_shade(t->params.ptr, frame, d1, d2, queue, counts, events, jobs, t->W, t->H, t->lut_lds,
                                         t->lds_bytes, lds_bits, shade_blocks, primary_blocks(*t), t->tile_shift, s));
        if (w != 0.0f) {
            if (w != cur_w) {
                HIP_TRY(erd_launch_adj_filter_t(dldr, adj, t->W, t->H, w, s));
                cur_w = w;
            }
            HIP_TRY(erd_launch_adj_wavefront(t->params.ptr, counts, jobs, events, queue, adj, tvis, frame, list, lcount, s1s,
                                             s2s, partials, blocks, (int)n, t->lds_bytes, s));
            HIP_TRY(erd_launch_adj_reduce(partials, 3 * blocks, out, s));
        }
        HIP_TRY(erd_launch_adj_reset(counts, lcount, s));
    }
    HIP_TRY(hipMemcpyAsync(dL_dlut, out, per_block * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ER_OK;
}

ER_API int er_adjoint_value(int tracer_id, const uint32_t* seeds1, const uint32_t* seeds2, size_t count, int passes,
                            const float* eval_lut, float* running_xyz)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    if ((st = adjoint_setup(*t, seeds1, seeds2, count, passes))) return st;
    if (!eval_lut || !running_xyz) return set_error(ER_ERROR, "pointer is NULL");
    const size_t n = (size_t)t->W * t->H;
    const DevParams& P = t->host_params;
    std::vector<float> sig((size_t)P.lut_max + 1);
    for (int i = 0; i <= P.lut_max; ++i)
        sig[i] = P.density_scale * erd::tf_lut_eval(eval_lut, t->tf_scale, (float)i, 0);
    DevBuf<uint32_t> d1, d2, qc;
    DevBuf<float> lut_e, sig_e;
    DevBuf<float4> frame, running;
    DevBuf<uchar4> display;
    if ((st = d1.alloc(n)) || (st = d2.alloc(n)) || (st = qc.alloc(4)) || (st = lut_e.alloc(ER_TF_LUT_NODES * 4)) ||
        (st = sig_e.alloc(sig.size())) || (st = frame.alloc(n)) || (st = running.alloc(n)) || (st = display.alloc(n)))
        return st;
    HIP_TRY(hipMemcpy(d1.ptr, seeds1, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d2.ptr, seeds2, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(lut_e.ptr, eval_lut, ER_TF_LUT_NODES * 4 * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(sig_e.ptr, sig.data(), sig.size() * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(running.ptr, 0, n * sizeof(float4)));
    /* value mode reads every sample: all occupancy bits set, plus the decision table */
    const size_t lds = (size_t)P.mc_words * 4 + ((size_t)P.lut_max + 1) * sizeof(float);
    for (int k = 0; k < passes; ++k) {
        HIP_TRY(erd_launch_adjoint(t->params.ptr, 1, lut_e.ptr, sig_e.ptr, nullptr, frame.ptr, nullptr, d1.ptr, d2.ptr,
                                   t->W, t->H, lds, t->stream));
        HIP_TRY(erd_launch_post(t->params.ptr, frame.ptr, running.ptr, display.ptr, k, qc.ptr, t->W, t->H, t->stream));
    }
    std::vector<float4> host(n);
    HIP_TRY(hipMemcpyAsync(host.data(), running.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    for (size_t i = 0; i < n; ++i) {
        running_xyz[3 * i] = host[i].x;
        running_xyz[3 * i + 1] = host[i].y;
        running_xyz[3 * i + 2] = host[i].z;
    }
    return ER_OK;
}

ER_API int er_get_estimate(int tracer_id, uint8_t* rgba, size_t size)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    const void* src = use_bilateral(*t) ? (const void*)t->disp_filt.ptr : (const void*)t->display.ptr;
    return copy_out(rgba, src, (size_t)t->W * t->H * 4, size, t->stream);
}

ER_API int er_get_frame_xyza(int tracer_id, float* xyza, size_t size)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    return copy_out(xyza, t->set[t->cur].frame.ptr, (size_t)t->W * t->H * 16, size, t->stream);
}

ER_API int er_get_running_xyza(int tracer_id, float* xyza, size_t size)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    return copy_out(xyza, t->running.ptr, (size_t)t->W * t->H * 16, size, t->stream);
}

ER_API int er_get_no_iterations(int tracer_id, int* n)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    if (!n) return set_error(ER_ERROR, "output pointer is NULL");
    *n = t->desc.no_iterations;
    return ER_OK;
}

ER_API int er_autofocus_seeds(int tracer_id, int film_u, int film_v, uint32_t* s0, uint32_t* s1)
{
    if (!s0 || !s1) return set_error(ER_ERROR, "output pointer is NULL");
    uint64_t z = 0x4175746f466f6375ull ^ ((uint64_t)(uint32_t)tracer_id << 32) ^ ((uint64_t)(uint32_t)film_v << 16) ^
                 (uint64_t)(uint32_t)film_u;
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    *s0 = (uint32_t)z ? (uint32_t)z : 1u;
    *s1 = (uint32_t)(z >> 32) ? (uint32_t)(z >> 32) : 1u;
    return ER_OK;
}

ER_API int er_get_autofocus_distance(int tracer_id, int film_u, int film_v, float* distance)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    if (!distance) return set_error(ER_ERROR, "output pointer is NULL");
    /* ComputeAutoFocusDistance (autofocus.cuh:26-75), which the reference entry point leaves
     * commented out (core.cu:145-148). The reference seeds it with rand(); here the stream is
     * a fixed function of the tracer ID and the film point, so the result is reproducible. */
    if ((st = build_params(*t))) return st;
    uint32_t s0, s1;
    er_autofocus_seeds(tracer_id, film_u, film_v, &s0, &s1);
    DevBuf<float> d;
    if ((st = d.alloc(1))) return st;
    HIP_TRY(erd_launch_autofocus(t->params.ptr, film_u, film_v, s0, s1, d.ptr, t->stream));
    HIP_TRY(hipMemcpyAsync(distance, d.ptr, sizeof(float), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    return ER_OK;
}

ER_API int er_get_seeds(int tracer_id, uint32_t* s1, uint32_t* s2, size_t count)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    const size_t n = (size_t)t->W * t->H;
    if ((st = copy_out(s1, t->seeds1.ptr, n * 4, count * 4, t->stream))) return st;
    return copy_out(s2, t->seeds2.ptr, n * 4, count * 4, t->stream);
}

ER_API int er_set_seeds(int tracer_id, const uint32_t* s1, const uint32_t* s2, size_t count)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    const size_t n = (size_t)t->W * t->H;
    if (!s1 || !s2) return set_error(ER_ERROR, "seed pointer is NULL");
    if (count < n) return set_error(ER_ERROR, "need %zu seeds, got %zu", n, count);
    HIP_TRY(hipMemcpy(t->seeds1.ptr, s1, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t->seeds2.ptr, s2, n * 4, hipMemcpyHostToDevice));
    return ER_OK;
}

ER_API int er_seed_tracer(int tracer_id, uint64_t seed)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    return tracer_seed(*t, seed);
}

ER_API int er_set_counting(int tracer_id, int on)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    t->counting = on ? 1 : 0;
    return ER_OK;
}

ER_API int er_get_counters(int tracer_id, er_counters* c)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    unsigned long long v[12];
    if ((st = copy_out(v, t->counters.ptr, sizeof(v), sizeof(v), t->stream))) return st;
    c->paths = v[0]; c->primary_samples = v[1]; c->shadow_samples = v[2]; c->shading_samples = v[3];
    c->primary_fetches = v[4]; c->shadow_fetches = v[5]; c->scatter_events = v[6]; c->shadow_rays = v[7];
    c->march_iterations = v[8]; c->skipped_samples = v[9]; c->skips_undone = v[10]; c->reserved = 0;
    return ER_OK;
}

ER_API int er_reset_counters(int tracer_id)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    HIP_TRY(hipMemsetAsync(t->counters.ptr, 0, 12 * sizeof(unsigned long long), t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    return ER_OK;
}

ER_API int er_get_gradient_magnitude_max(int tracer_id, float* gm)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    int st;
    Tracer* t = get_tracer(tracer_id, st);
    if (!t) return st;
    if ((st = build_params(*t))) return st;
    *gm = t->gm_max;
    return ER_OK;
}

ER_API int er_neglog_table_entry(uint32_t m, float* value)
{
    std::lock_guard<std::recursive_mutex> lk(g_mutex);
    if (!value) return set_error(ER_ERROR, "output pointer is NULL");
    if (m >= (1u << ER_NEGLOG_BITS)) return set_error(ER_ERROR, "m out of range");
    build_neglog_host();
    *value = g_neglog_host[m];
    return ER_OK;
}

} /* extern "C" */
