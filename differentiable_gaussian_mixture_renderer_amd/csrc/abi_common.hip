// SPDX-License-Identifier: GPL-3.0-or-later
// Error reporting, camera set-up and workspace sizing shared by the C ABI.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "host_common.h"
#include "primitives.h"

namespace dgmr {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

bool valid_tiling(const dgmr_tiling* t)
{
    if (!t)
        return false;
    const bool shape = (t->tile_w == 16 && t->tile_h == 16) || (t->tile_w == 8 && t->tile_h == 8) || (t->tile_w == 32 && t->tile_h == 8);
    const bool batch = t->batch == 64 || t->batch == 128 || t->batch == 256;
    return shape && batch && (t->cull == 0 || t->cull == 1) && (t->slot_order == 0 || t->slot_order == 1);
}

uint32_t higher_msb(uint32_t n)
{
    uint32_t r = 0;
    while (r < 32 && (n >> r) != 0)
        ++r;
    return r;
}

// 4x4 inverse in double (Gauss-Jordan with partial pivoting)
static bool invert4(const double* m, double* out)
{
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 8; ++c)
            a[r][c] = c < 4 ? m[4 * r + c] : (c - 4 == r ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c]))
                piv = r;
        if (a[piv][c] == 0)
            return false;
        if (piv != c)
            for (int k = 0; k < 8; ++k)
                std::swap(a[piv][k], a[c][k]);
        const double inv = 1.0 / a[c][c];
        for (int k = 0; k < 8; ++k)
            a[c][k] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == c)
                continue;
            const double f = a[r][c];
            for (int k = 0; k < 8; ++k)
                a[r][k] -= f * a[c][k];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c)
            out[4 * r + c] = a[r][c + 4];
    return true;
}

template <class S>
DevCamera<S> make_dev_camera(const dgmr_camera& c, const dgmr_tiling& t)
{
    DevCamera<S> d;
    // matrices arrive as the scene dtype would hold them
    double proj_s[16];
    for (int i = 0; i < 16; ++i) {
        d.view[i] = S(c.view[i]);
        d.proj[i] = S(c.proj[i]);
        proj_s[i] = double(d.proj[i]);
    }
    double inv[16];
    if (!invert4(proj_s, inv))
        std::memset(inv, 0, sizeof(inv));
    for (int i = 0; i < 16; ++i)
        d.inv_proj[i] = S(inv[i]);
    for (int i = 0; i < 3; ++i) {
        d.pos[i] = S(c.position[i]);
        d.bg[i] = S(c.background[i]);
    }
    // the reference passes tan_fov as float even for double scenes (vol_marcher.cu:229-230)
    d.tan_fovx = S(float(c.tan_fovx));
    d.tan_fovy = S(float(c.tan_fovy));
    d.focal_y = S(c.height) / (S(2) * d.tan_fovy);
    d.focal_x = S(c.width) / (S(2) * d.tan_fovx);
    d.width = c.width;
    d.height = c.height;
    d.tile_w = t.tile_w;
    d.tile_h = t.tile_h;
    d.grid_x = (c.width + t.tile_w - 1) / t.tile_w;
    d.grid_y = (c.height + t.tile_h - 1) / t.tile_h;
    return d;
}

template DevCamera<float> make_dev_camera<float>(const dgmr_camera&, const dgmr_tiling&);

// ---------------------------------------------------------------- stage timing
namespace {
struct Record {
    const char* name;
    hipEvent_t start, stop;
    double bytes;
};
struct Profiler {
    bool on = false;
    std::vector<Record> records;
    std::vector<hipEvent_t> pool;
    hipEvent_t get()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess)
            return nullptr;
        return e;
    }
};
std::mutex g_prof_mutex;
Profiler g_prof;
} // namespace

StageScope::StageScope(const char* name, hipStream_t s, double bytes) : stream(s)
{
    std::lock_guard<std::mutex> lk(g_prof_mutex);
    if (!g_prof.on)
        return;
    Record r { name, g_prof.get(), g_prof.get(), bytes };
    if (!r.start || !r.stop)
        return;
    (void)hipEventRecord(r.start, s);
    g_prof.records.push_back(r);
    slot = int(g_prof.records.size()) - 1;
}

void stage_set_bytes(int slot, double bytes)
{
    if (slot < 0)
        return;
    std::lock_guard<std::mutex> lk(g_prof_mutex);
    if (slot < int(g_prof.records.size()))
        g_prof.records[slot].bytes = bytes;
}

// The pair-count read-back slots: one pinned host word and one event per
// calling thread.  A thread takes a slot on its first count and hands it back
// to a process-wide free list when it ends (no HIP call at thread exit), so
// the pinned memory is bounded by the largest number of threads that have
// counted at the same time, however many threads come and go.
namespace {
struct CountSlot {
    uint32_t* host = nullptr;
    hipEvent_t done = nullptr;
};
std::mutex g_slot_mutex;
std::vector<CountSlot>* g_free_slots = new std::vector<CountSlot>(); // never destroyed: threads may end after static destructors
struct ThreadSlot {
    CountSlot slot;
    ~ThreadSlot()
    {
        if (!slot.host)
            return;
        std::lock_guard<std::mutex> lk(g_slot_mutex);
        g_free_slots->push_back(slot);
    }
};
thread_local ThreadSlot t_slot;

int acquire_count_slot(CountSlot* out)
{
    {
        std::lock_guard<std::mutex> lk(g_slot_mutex);
        if (!g_free_slots->empty()) {
            *out = g_free_slots->back();
            g_free_slots->pop_back();
            return DGMR_OK;
        }
    }
    void* p = nullptr;
    DGMR_HIP_CHECK(hipHostMalloc(&p, 256, hipHostMallocDefault));
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipHostFree(p);
        set_error("hipEventCreateWithFlags failed for the pair count");
        return DGMR_ERR_DEVICE;
    }
    out->host = static_cast<uint32_t*>(p);
    out->done = e;
    return DGMR_OK;
}
} // namespace

int enqueue_count(const uint32_t* dev_count, hipStream_t st, PendingCount* pc)
{
    if (!t_slot.slot.host) {
        const int rc = acquire_count_slot(&t_slot.slot);
        if (rc != DGMR_OK)
            return rc;
    }
    uint32_t* host = t_slot.slot.host;
    hipEvent_t done = t_slot.slot.done;
    DGMR_HIP_CHECK(hipMemcpyAsync(host, dev_count, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    DGMR_HIP_CHECK(hipEventRecord(done, st));
    pc->device = dev_count;
    pc->host = host;
    pc->done = done;
    return DGMR_OK;
}

int64_t PendingCount::wait()
{
    if (!done || !host)
        return 0;
    if (hipEventSynchronize(done) != hipSuccess) {
        set_error("waiting for the pair count failed");
        return -1;
    }
    const int64_t r = int64_t(*host);
    for (int k = 0; k < n_slots; ++k)
        stage_set_bytes(slot[k], per_pair[k] * double(r) + fixed[k]);
    return r;
}

StageScope::~StageScope()
{
    if (slot < 0)
        return;
    std::lock_guard<std::mutex> lk(g_prof_mutex);
    if (slot < int(g_prof.records.size()))
        (void)hipEventRecord(g_prof.records[slot].stop, stream);
}
template DevCamera<double> make_dev_camera<double>(const dgmr_camera&, const dgmr_tiling&);

} // namespace dgmr

extern "C" {

const char* dgmr_last_error(void) { return dgmr::g_last_error.c_str(); }

const char* dgmr_version(void) { return "dgmr-mi355x 0.1 (gfx950)"; }

int dgmr_profile_enable(int on)
{
    std::lock_guard<std::mutex> lk(dgmr::g_prof_mutex);
    dgmr::g_prof.on = on != 0;
    return DGMR_OK;
}

int dgmr_profile_read(dgmr_stage_time* out, int32_t max_stages, int32_t* n_stages)
{
    std::lock_guard<std::mutex> lk(dgmr::g_prof_mutex);
    std::vector<dgmr_stage_time> acc;
    for (auto& r : dgmr::g_prof.records) {
        float ms = 0.f;
        if (hipEventSynchronize(r.stop) != hipSuccess || hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess)
            ms = 0.f;
        size_t k = 0;
        while (k < acc.size() && std::strcmp(acc[k].name, r.name) != 0)
            ++k;
        if (k == acc.size())
            acc.push_back(dgmr_stage_time { r.name, 0.0, 0, 0.0 });
        acc[k].total_ms += ms;
        acc[k].launches += 1;
        acc[k].bytes += r.bytes;
        dgmr::g_prof.pool.push_back(r.start);
        dgmr::g_prof.pool.push_back(r.stop);
    }
    dgmr::g_prof.records.clear();
    const int32_t count = int32_t(acc.size());
    if (n_stages)
        *n_stages = count;
    for (int32_t i = 0; i < count && i < max_stages && out; ++i)
        out[i] = acc[i];
    return DGMR_OK;
}

} // extern "C"
