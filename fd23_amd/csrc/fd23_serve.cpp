// Request-serving engine: continuous batching of fd23 solves on one GPU.
//
// One engine thread owns every device call.  Its loop:
//   1. admission  - take queued requests in arrival order; merge the ones
//                   that may share a context (same model key and physics);
//                   admit a batch when a stream slot is idle and its device
//                   bytes fit in the budget left by resident contexts
//                   (smaller requests may pass one that does not fit);
//   2. scheduling - give every resident context its next chunk of steps
//                   (graph replays through fd23_run) while it has fewer
//                   than two chunks queued on its stream;
//   3. retirement - when a context's last chunk has completed, copy each
//                   request's seismograms to host memory, destroy the
//                   context (its HBM returns to the budget) and wake the
//                   waiters.
// Client threads only touch host-side request records under the mutex.
#include "fd23_serve.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

extern "C" void fd23_set_error_(const char *msg);

namespace {

using Clock = std::chrono::steady_clock;

double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Request {
    long long id = 0;
    fd23_request r{};             // scalar fields; pointers re-aimed at the copies below
    std::vector<std::vector<char>> mat; // material arrays, or vp vs den with from_model
    std::vector<const void *> matp;
    bool from_model = false;
    std::vector<double> cpml, stf, M;
    std::vector<int> xs, zs, rx, rz;
    int ncomp = 1, nframes = 0, storage = 0;
    size_t csize = 4;
    Clock::time_point t_submit, t_admit, t_done;
    int state = FD23_REQ_QUEUED;
    int err = FD23_OK;
    std::string msg;
    std::vector<char> seis;       // [shot][comp][frame][rec]
};

// A request inside a context: its shots occupy some of the context's shot
// slots and start at context step t0 (fd23_set_shot_slot), so its
// seismograms read back from frame 0 as if it had run alone.
struct Member {
    std::shared_ptr<Request> q;
    std::vector<int> slots; // the context's shot slot of each request shot
    int t0 = 0;
    bool done = false;
    int end() const { return t0 + q->r.nt; }
};

struct Group {
    fd23_ctx *ctx = nullptr;
    int slot = -1;                  // stream slot
    std::vector<Member> mem;
    std::vector<int> owner;         // per shot slot: member index, -1 idle
    int nshots = 0, itrecord = 1, max_frames = 0;
    int next_it = 0;                // context steps enqueued
    int done_it = 0;                // context steps completed
    size_t bytes = 0;
    std::deque<std::pair<hipEvent_t, int>> inflight; // chunk event, the step it reaches
    bool failed = false;
    bool box = false;
    // the end of the next member to finish after step it, or -1
    int next_end(int it) const {
        int e = -1;
        for (const auto &m : mem)
            if (!m.done && m.end() > it && (e < 0 || m.end() < e)) e = m.end();
        return e;
    }
    int free_slots() const { return (int)std::count(owner.begin(), owner.end(), -1); }
};

int material_count(int wave) { return wave == FD23_WAVE_PSV ? 5 : 3; }

// requests that may share one context
bool compatible(const Request &a, const Request &b) {
    const fd23_request &x = a.r, &y = b.r;
    if (x.model_key == 0 || x.model_key != y.model_key) return false;
    // what the context holds: the model, grid, time step, C-PML, storage,
    // receiver capacity and box.  Sources are set per shot, and the number
    // of steps may differ: each request's shots keep their own first step
    // and its results are read out when its last step is done.
    if (std::make_tuple(x.wave, a.storage, x.nx, x.nz, x.npml, x.topabsorb, x.itrecord, x.nrec, x.box) !=
        std::make_tuple(y.wave, b.storage, y.nx, y.nz, y.npml, y.topabsorb, y.itrecord, y.nrec, y.box))
        return false;
    if (x.dx != y.dx || x.dt != y.dt || a.cpml != b.cpml) return false;
    if (x.box &&
        (std::make_tuple(x.box_x0, x.box_z0, x.box_bx, x.box_bz, x.box_incident) !=
             std::make_tuple(y.box_x0, y.box_z0, y.box_bx, y.box_bz, y.box_incident) ||
         x.box_angle != y.box_angle || x.box_f0 != y.box_f0 || x.box_t0 != y.box_t0 || x.box_amp != y.box_amp ||
         x.box_vp != y.box_vp || x.box_vs != y.box_vs || x.box_rho != y.box_rho))
        return false;
    return true;
}

fd23_desc make_desc(const Request &q, int nshots, void *stream) {
    fd23_desc d{};
    d.wave = q.r.wave;
    d.storage = q.storage;
    d.nx = q.r.nx;
    d.nz = q.r.nz;
    d.npml = q.r.npml;
    d.topabsorb = q.r.topabsorb;
    d.nshots = nshots;
    d.max_frames = q.nframes;
    d.max_rec = q.r.nrec;
    d.stream = stream;
    return d;
}

size_t batch_bytes(const Request &q, int nshots, int frames) {
    fd23_desc d = make_desc(q, nshots, nullptr);
    d.max_frames = frames;
    size_t b = fd23_desc_device_bytes(&d);
    if (q.r.box) b += (size_t)nshots * 5 * q.r.box_bx * q.r.box_bz * q.csize + 256;
    return b + ((size_t)1 << 20); // graphs, event lists, tables
}

} // namespace

struct fd23_server {
    fd23_server_desc d{};
    int device = 0;
    std::vector<hipStream_t> streams;
    std::vector<bool> slot_busy;
    std::thread engine;
    std::mutex mu;
    std::condition_variable cv_engine, cv_done;
    bool stop = false;
    long long next_id = 1;
    std::deque<std::shared_ptr<Request>> queue;
    std::map<long long, std::shared_ptr<Request>> all;
    std::vector<std::unique_ptr<Group>> groups; // engine thread only
    size_t used = 0;
    // tuned tile per shape: wave, storage, nx, nz, shots, npml, topabsorb and
    // box, since a box or a first-arrival window runs some shapes as others
    std::map<std::tuple<int, int, int, int, int, int, int, int>, int> tiles;
    fd23_server_stats st{};

    void run();
    bool admit();
    bool schedule();
    bool retire();
    int start_group(Group &g);
    int place(Group &g, Member &m); // sources and receivers of a member's slots
    // shot slots of a new context for a batch of `shots` shots: spares for
    // requests that may join (shared problem, no box), within the 64 slots
    // the kernels can leave idle
    int capacity(const Request &q, int shots) const {
        if (d.spare_shots <= 0 || q.r.model_key == 0 || q.r.box || shots >= 64) return shots;
        return std::min(shots + d.spare_shots, 64);
    }
    int join(Group &g, const std::shared_ptr<Request> &q, const std::vector<int> &slots);
    int harvest(Group &g, Member &m);
    int quiet(Group &g, int shot);
    void finish_group(Group &g);
    void complete(Request &q, bool failed_ctx);
};

int fd23_server::place(Group &g, Member &m) {
    const Request &q = *m.q;
    int r;
    for (int s = 0; s < q.r.nshots; s++) {
        const int shot = m.slots[s];
        if (q.r.src_kind >= 0 &&
            (r = fd23_set_source(g.ctx, shot, q.r.src_kind, q.xs[s], q.zs[s], q.stf.data(), q.r.lsrc, q.M.data())))
            return r;
        for (int c = 0; c < q.ncomp; c++)
            if (q.r.nrec > 0 && (r = fd23_set_receivers(g.ctx, shot, c, q.rx.data(), q.rz.data(), q.r.nrec)))
                return r;
    }
    return FD23_OK;
}

int fd23_server::start_group(Group &g) {
    const Request &q0 = *g.mem[0].q;
    fd23_desc d = make_desc(q0, g.nshots, streams[g.slot]);
    d.max_frames = g.max_frames; // the longest member
    int r;
    if ((r = fd23_ctx_create(&d, &g.ctx))) return r;
    if (q0.from_model) {
        double stab;
        r = fd23_set_model(g.ctx, q0.matp[0], q0.matp[1], q0.matp[2], q0.r.dx, q0.r.dt, &stab);
    } else {
        r = fd23_set_material(g.ctx, q0.matp.data());
    }
    if (r) return r;
    if ((r = fd23_set_cpml(g.ctx, q0.cpml.data()))) return r;
    for (auto &m : g.mem)
        if ((r = place(g, m))) return r;
    for (int s = 0; s < g.nshots; s++)
        if (g.owner[s] == -1 && (r = fd23_set_shot_slot(g.ctx, s, 0, 0))) return r;
    const fd23_request &b = q0.r;
    if (b.box && (r = fd23_box_configure(g.ctx, -1, b.box_x0, b.box_z0, b.box_bx, b.box_bz, b.box_incident,
                                         b.box_angle, b.box_f0, b.box_t0, b.box_amp, b.box_vp, b.box_vs,
                                         b.box_rho, b.dx, b.dt)))
        return r;
    if (this->d.autotune) {
        const auto key = std::make_tuple(b.wave, q0.storage, b.nx, b.nz, g.nshots, b.npml, b.topabsorb, b.box ? 1 : 0);
        auto it = tiles.find(key);
        if (it == tiles.end()) {
            double ms = 0;
            if ((r = fd23_autotune(g.ctx, 4, &ms))) return r;
            tiles[key] = fd23_get_tile(g.ctx);
        } else if ((r = fd23_set_tile(g.ctx, it->second))) {
            return r;
        }
    }
    return FD23_OK;
}

// admission: returns true if anything was admitted.  A queued request first
// joins a running context that holds a compatible request and enough idle
// shot slots (its shots start at that context's next step); the rest are
// merged into new contexts as HBM and stream slots allow.
bool fd23_server::admit() {
    std::vector<std::vector<std::shared_ptr<Request>>> batches;
    std::vector<std::tuple<Group *, std::shared_ptr<Request>, std::vector<int>>> joins;
    {
        std::lock_guard<std::mutex> lk(mu);
        size_t free_dev = 0, total_dev = 0;
        if (!queue.empty() && hipMemGetInfo(&free_dev, &total_dev) != hipSuccess) free_dev = 0;
        int idle = 0;
        for (bool b : slot_busy) idle += !b;
        std::vector<bool> taken(queue.size(), false);
        // joins into running contexts (slots reserved here, placed below)
        for (size_t i = 0; i < queue.size(); i++) {
            const Request &q = *queue[i];
            for (auto &gp : groups) {
                Group &g = *gp;
                if (g.failed || g.box || q.r.box || g.mem.empty() || !compatible(*g.mem[0].q, q) ||
                    q.nframes > g.max_frames || g.free_slots() < q.r.nshots || g.next_end(g.next_it) < 0)
                    continue;
                std::vector<int> slots;
                for (int s = 0; s < g.nshots && (int)slots.size() < q.r.nshots; s++)
                    if (g.owner[s] == -1) {
                        g.owner[s] = -2; // reserved
                        slots.push_back(s);
                    }
                joins.emplace_back(&g, queue[i], slots);
                taken[i] = true;
                break;
            }
        }
        size_t planned = 0;
        for (size_t i = 0; i < queue.size() && idle > 0; i++) {
            if (taken[i]) continue;
            std::vector<size_t> pick{i};
            int shots = queue[i]->r.nshots;
            for (size_t j = i + 1; j < queue.size(); j++) {
                if (taken[j] || !compatible(*queue[i], *queue[j])) continue;
                if (shots + queue[j]->r.nshots > d.max_batch_shots) continue;
                pick.push_back(j);
                shots += queue[j]->r.nshots;
            }
            int frames = 0;
            for (size_t k : pick) frames = std::max(frames, queue[k]->nframes);
            const size_t need = batch_bytes(*queue[i], capacity(*queue[i], shots), frames);
            if (pick.size() == 1 && need > d.hbm_budget) {
                // can never fit: fail it rather than let it wait forever
                taken[i] = true;
                queue[i]->state = FD23_REQ_FAILED;
                queue[i]->err = FD23_ENOMEM;
                queue[i]->msg = "request needs more device memory than the server budget";
                queue[i]->t_done = Clock::now();
                st.failed++;
                cv_done.notify_all();
                continue;
            }
            const size_t left = d.hbm_budget > used + planned ? d.hbm_budget - used - planned : 0;
            const size_t dev_left = free_dev > planned + (size_t(256) << 20) ? free_dev - planned - (size_t(256) << 20) : 0;
            if (need > left || need > dev_left) continue; // waits; smaller ones may pass
            planned += need;
            idle--;
            std::vector<std::shared_ptr<Request>> b;
            for (size_t k : pick) {
                taken[k] = true;
                b.push_back(queue[k]);
            }
            batches.push_back(std::move(b));
        }
        std::deque<std::shared_ptr<Request>> rest;
        for (size_t i = 0; i < queue.size(); i++)
            if (!taken[i]) rest.push_back(queue[i]);
        queue.swap(rest);
        if (batches.empty() && joins.empty()) return false;
        const auto now = Clock::now();
        for (auto &b : batches)
            for (auto &q : b) {
                q->state = FD23_REQ_RUNNING;
                q->t_admit = now;
            }
        for (auto &j : joins) {
            std::get<1>(j)->state = FD23_REQ_RUNNING;
            std::get<1>(j)->t_admit = now;
        }
    }
    for (auto &j : joins) {
        Group &g = *std::get<0>(j);
        const int r = join(g, std::get<1>(j), std::get<2>(j));
        std::lock_guard<std::mutex> lk(mu);
        if (r) {
            const char *m = fd23_last_error();
            Request &q = *std::get<1>(j);
            q.err = r;
            q.msg = m ? m : "";
            q.state = FD23_REQ_FAILED;
            q.t_done = Clock::now();
            st.failed++;
            cv_done.notify_all();
        } else {
            st.joined++;
        }
    }
    for (auto &b : batches) {
        auto g = std::make_unique<Group>();
        int shot = 0;
        for (auto &q : b) {
            Member m;
            m.q = q;
            for (int s = 0; s < q->r.nshots; s++) m.slots.push_back(shot++);
            g->mem.push_back(std::move(m));
            g->max_frames = std::max(g->max_frames, q->nframes);
        }
        g->nshots = capacity(*b[0], shot); // spare slots idle until a request joins
        g->owner.assign(g->nshots, -1);
        for (size_t k = 0; k < g->mem.size(); k++)
            for (int s : g->mem[k].slots) g->owner[s] = (int)k;
        g->itrecord = b[0]->r.itrecord;
        g->box = b[0]->r.box != 0;
        g->bytes = batch_bytes(*b[0], g->nshots, g->max_frames);
        for (size_t s = 0; s < slot_busy.size(); s++)
            if (!slot_busy[s]) {
                g->slot = (int)s;
                slot_busy[s] = true;
                break;
            }
        used += g->bytes;
        int r = start_group(*g);
        {
            std::lock_guard<std::mutex> lk(mu);
            st.launches++;
            if (b.size() > 1) st.merged += (long long)b.size();
            if (r) {
                const char *m = fd23_last_error();
                for (auto &q : b) {
                    q->err = r;
                    q->msg = m ? m : "";
                }
            }
        }
        if (r) g->failed = true;
        groups.push_back(std::move(g));
    }
    std::lock_guard<std::mutex> lk(mu);
    st.peak_active = std::max<long long>(st.peak_active, (long long)groups.size());
    return true;
}

// a request enters idle slots of a running context: its shots start at the
// context's next step, from zero fields
int fd23_server::join(Group &g, const std::shared_ptr<Request> &q, const std::vector<int> &slots) {
    // the chunks in flight finish first: the slots' arrays and the event
    // lists change under them
    if (hipStreamSynchronize(streams[g.slot]) != hipSuccess) return FD23_EHIP;
    Member m;
    m.q = q;
    m.slots = slots;
    m.t0 = g.next_it;
    int r = FD23_OK;
    for (int s : slots)
        if ((r = fd23_reset_shot(g.ctx, s)) || (r = fd23_set_shot_slot(g.ctx, s, 1, m.t0))) break;
    if (!r) r = place(g, m);
    if (r) {
        for (int s : slots) {
            g.owner[s] = -1;
            (void)fd23_set_shot_slot(g.ctx, s, 0, 0);
        }
        return r;
    }
    for (int s : slots) g.owner[s] = (int)g.mem.size();
    g.mem.push_back(std::move(m));
    return FD23_OK;
}

// give every resident context its next chunk; true if anything was queued.
// Chunks end where a request ends, so that its results are final when the
// chunk's event completes.
bool fd23_server::schedule() {
    bool any = false;
    for (auto &gp : groups) {
        Group &g = *gp;
        while (!g.failed && g.inflight.size() < 2) {
            const int e = g.next_end(g.next_it);
            if (e < 0) break;
            const int n = std::min(d.chunk_steps, e - g.next_it);
            int r = fd23_run(g.ctx, g.next_it, n, g.itrecord);
            hipEvent_t ev = nullptr;
            if (!r && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) r = FD23_EHIP;
            if (!r && hipEventRecord(ev, streams[g.slot]) != hipSuccess) r = FD23_EHIP;
            if (r) {
                if (ev) (void)hipEventDestroy(ev);
                const char *m = fd23_last_error();
                std::lock_guard<std::mutex> lk(mu);
                for (auto &mb : g.mem) {
                    mb.q->err = r;
                    mb.q->msg = m && *m ? m : "chunk launch failed";
                }
                g.failed = true;
                break;
            }
            g.next_it += n;
            g.inflight.emplace_back(ev, g.next_it);
            any = true;
        }
    }
    return any;
}

// an idle slot: zero state and no source samples, so that it stays zero on
// the paths that still step idle slots (shared FP16/BF16 scales never see
// it), and skipped where the kernels can
int fd23_server::quiet(Group &g, int shot) {
    int r;
    if ((r = fd23_reset_shot(g.ctx, shot))) return r;
    if ((r = fd23_set_source(g.ctx, shot, FD23_SRC_LINE, 0, 0, nullptr, 0, nullptr))) return r;
    // slots from 64 on cannot be idle: they step zeros, which stay zero
    return fd23_set_shot_slot(g.ctx, shot, shot < 64 ? 0 : 1, 0);
}

// copy a finished member's seismograms out
int fd23_server::harvest(Group &g, Member &m) {
    Request &q = *m.q;
    const size_t per = (size_t)q.nframes * q.r.nrec * q.csize;
    q.seis.assign(per * q.r.nshots * q.ncomp, 0);
    std::vector<char> tmp(per);
    for (int s = 0; s < q.r.nshots; s++)
        for (int c = 0; c < q.ncomp; c++) {
            if (q.r.nrec == 0) continue;
            int r = fd23_get_seismograms(g.ctx, m.slots[s], c, (float *)tmp.data(), q.nframes);
            if (r) return r;
            std::memcpy(q.seis.data() + ((size_t)s * q.ncomp + c) * per, tmp.data(), per);
        }
    return FD23_OK;
}

// state and statistics of a request leaving the engine (lock held)
void fd23_server::complete(Request &q, bool failed_ctx) {
    q.t_done = Clock::now();
    if (failed_ctx && !q.err) {
        q.err = FD23_EHIP;
        q.msg = "context failed";
    }
    q.state = q.err ? FD23_REQ_FAILED : FD23_REQ_DONE;
    if (q.err) {
        st.failed++;
    } else {
        st.completed++;
        st.cell_updates += (double)q.r.nshots * q.r.nx * q.r.nz * q.r.nt;
    }
}

void fd23_server::finish_group(Group &g) {
    // the stream is idle here (last event completed) unless the group failed
    if (g.ctx) (void)hipStreamSynchronize(streams[g.slot]);
    for (auto &m : g.mem) {
        if (m.done) continue;
        if (!g.failed && !m.q->err) {
            int r = harvest(g, m);
            if (r) {
                m.q->err = r;
                m.q->msg = fd23_last_error();
            }
        }
    }
    for (auto &e : g.inflight) (void)hipEventDestroy(e.first);
    g.inflight.clear();
    if (g.ctx) fd23_ctx_destroy(g.ctx);
    g.ctx = nullptr;
    used -= g.bytes;
    slot_busy[g.slot] = false;
    std::lock_guard<std::mutex> lk(mu);
    for (auto &m : g.mem)
        if (!m.done) {
            m.done = true;
            complete(*m.q, g.failed);
        }
    cv_done.notify_all();
}

// retire finished requests (their slots go idle) and finished contexts;
// true if any retired
bool fd23_server::retire() {
    bool any = false;
    for (size_t i = 0; i < groups.size();) {
        Group &g = *groups[i];
        while (!g.inflight.empty() && hipEventQuery(g.inflight.front().first) == hipSuccess) {
            g.done_it = g.inflight.front().second;
            (void)hipEventDestroy(g.inflight.front().first);
            g.inflight.pop_front();
        }
        bool open = false; // a member still runs
        if (!g.failed)
            for (size_t k = 0; k < g.mem.size(); k++) {
                Member &m = g.mem[k];
                if (m.done) continue;
                if (m.end() > g.done_it) {
                    open = true;
                    continue;
                }
                if (g.next_end(g.next_it) < 0 && g.inflight.empty()) continue; // the group ends: harvested there
                // a request ends inside a running context: results out, slots idle
                int r = harvest(g, m);
                for (int s : m.slots) {
                    g.owner[s] = -1;
                    // a box context runs every shot (and takes no joins)
                    if (!r && !g.box) r = quiet(g, s);
                }
                if (r) {
                    m.q->err = r;
                    m.q->msg = fd23_last_error();
                }
                m.done = true;
                any = true;
                std::lock_guard<std::mutex> lk(mu);
                complete(*m.q, false);
                cv_done.notify_all();
            }
        if (g.failed || (!open && g.inflight.empty())) {
            finish_group(g);
            groups.erase(groups.begin() + i);
            any = true;
        } else {
            i++;
        }
    }
    return any;
}

void fd23_server::run() {
    (void)hipSetDevice(device);
    for (;;) {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (stop && groups.empty()) break;
        }
        bool busy = false;
        busy |= retire();
        {
            bool stopping;
            {
                std::lock_guard<std::mutex> lk(mu);
                stopping = stop;
            }
            if (!stopping) busy |= admit();
        }
        busy |= schedule();
        if (!busy) {
            std::unique_lock<std::mutex> lk(mu);
            // new submissions wake the engine at once; otherwise poll the
            // in-flight chunks again shortly
            cv_engine.wait_for(lk, std::chrono::microseconds(groups.empty() ? 20000 : 200));
        }
    }
}

extern "C" int fd23_server_create(const fd23_server_desc *desc, fd23_server **out) {
    if (!desc || !out) {
        fd23_set_error_("null argument");
        return FD23_EINVAL;
    }
    if (desc->max_active < 1 || desc->max_active > 64 || desc->chunk_steps < 1 || desc->max_batch_shots < 1 ||
        desc->spare_shots < 0) {
        fd23_set_error_("bad server description");
        return FD23_EINVAL;
    }
    auto s = std::make_unique<fd23_server>();
    s->d = *desc;
    if (hipGetDevice(&s->device) != hipSuccess) {
        fd23_set_error_("no HIP device");
        return FD23_EHIP;
    }
    if (s->d.hbm_budget == 0) {
        size_t f = 0, t = 0;
        if (hipMemGetInfo(&f, &t) != hipSuccess) {
            fd23_set_error_("hipMemGetInfo failed");
            return FD23_EHIP;
        }
        s->d.hbm_budget = (size_t)(0.9 * (double)f);
    }
    s->streams.resize(s->d.max_active);
    for (auto &st : s->streams)
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            fd23_set_error_("stream creation failed");
            for (auto &x : s->streams)
                if (x) (void)hipStreamDestroy(x);
            return FD23_EHIP;
        }
    s->slot_busy.assign(s->d.max_active, false);
    s->st.budget = s->d.hbm_budget;
    fd23_server *p = s.release();
    p->engine = std::thread([p] { p->run(); });
    *out = p;
    return FD23_OK;
}

extern "C" int fd23_server_destroy(fd23_server *s) {
    if (!s) return FD23_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (auto &q : s->queue) {
            q->state = FD23_REQ_FAILED;
            q->err = FD23_EINVAL;
            q->msg = "server stopped";
        }
        s->queue.clear();
        s->cv_engine.notify_all();
        s->cv_done.notify_all();
    }
    s->engine.join();
    for (auto &x : s->streams) (void)hipStreamDestroy(x);
    delete s;
    return FD23_OK;
}

extern "C" int fd23_server_submit(fd23_server *s, const fd23_request *r, long long *id) {
    if (!s || !r || !id) {
        fd23_set_error_("null argument");
        return FD23_EINVAL;
    }
    const bool psv = r->wave == FD23_WAVE_PSV;
    if ((r->wave != FD23_WAVE_PSV && r->wave != FD23_WAVE_SH) || r->storage < FD23_STORE_F32 ||
        r->storage > FD23_STORE_F64 || r->nshots < 1 || r->nt < 1 || r->itrecord < 1 || r->nrec < 0 ||
        (!r->material && !r->model) || !r->cpml || (r->src_kind >= 0 && (!r->xs || !r->zs || !r->M || r->lsrc < 0 ||
                                                          (r->lsrc > 0 && !r->stf))) ||
        (r->nrec > 0 && (!r->rx || !r->rz)) || (r->box && !psv)) {
        fd23_set_error_("bad request");
        return FD23_EINVAL;
    }
    auto q = std::make_shared<Request>();
    q->r = *r;
    q->storage = r->storage;
    if (r->box && (r->storage == FD23_STORE_F16 || r->storage == FD23_STORE_BF16)) q->storage = FD23_STORE_F32;
    q->ncomp = psv ? 2 : 1;
    q->nframes = (r->nt + r->itrecord - 1) / r->itrecord;
    q->csize = q->storage == FD23_STORE_F64 ? 8 : 4;
    const size_t n = (size_t)r->nx * r->nz;
    const size_t msize = r->storage == FD23_STORE_F64 ? 8 : 4;
    q->from_model = !r->material;
    if (q->from_model) { // vp vs den: SH leaves vp empty
        if ((!r->model[1] || !r->model[2]) || (psv && !r->model[0])) {
            fd23_set_error_("null model array");
            return FD23_EINVAL;
        }
        if (!psv && r->storage == FD23_STORE_F64) {
            fd23_set_error_("SH material from the model is computed in float: pass the material");
            return FD23_EINVAL;
        }
        for (int i = 0; i < 3; i++) {
            const char *p = (const char *)r->model[i];
            if (p)
                q->mat.emplace_back(p, p + n * msize);
            else
                q->mat.emplace_back();
        }
    } else {
        const int nm = material_count(r->wave);
        for (int i = 0; i < nm; i++) {
            if (!r->material[i]) {
                fd23_set_error_("null material array");
                return FD23_EINVAL;
            }
            const char *p = (const char *)r->material[i];
            q->mat.emplace_back(p, p + n * msize);
        }
    }
    if (msize == 8 && q->csize == 4) {
        fd23_set_error_("FP64 material needs FP64 storage");
        return FD23_EINVAL;
    }
    for (auto &m : q->mat) q->matp.push_back(m.empty() ? nullptr : m.data());
    q->cpml.assign(r->cpml, r->cpml + fd23_cpml_pack_len(r->wave, r->nx, r->nz));
    if (r->src_kind >= 0) {
        q->xs.assign(r->xs, r->xs + r->nshots);
        q->zs.assign(r->zs, r->zs + r->nshots);
        q->stf.assign(r->stf, r->stf + r->lsrc);
        q->M.assign(r->M, r->M + 3);
    }
    if (r->nrec > 0) {
        q->rx.assign(r->rx, r->rx + r->nrec);
        q->rz.assign(r->rz, r->rz + r->nrec);
    }
    q->r.material = nullptr;
    q->r.model = nullptr;
    q->r.cpml = nullptr;
    q->r.xs = q->r.zs = q->r.rx = q->r.rz = nullptr;
    q->r.stf = q->r.M = nullptr;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        fd23_set_error_("server stopped");
        return FD23_EINVAL;
    }
    q->id = s->next_id++;
    q->t_submit = Clock::now();
    s->queue.push_back(q);
    s->all[q->id] = q;
    s->st.submitted++;
    *id = q->id;
    s->cv_engine.notify_all();
    return FD23_OK;
}

extern "C" int fd23_server_wait(fd23_server *s, long long id, double timeout_s, int *state) {
    if (!s) {
        fd23_set_error_("null server");
        return FD23_EINVAL;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) {
        fd23_set_error_("unknown request id");
        return FD23_EINVAL;
    }
    auto q = it->second;
    auto done = [&] { return q->state == FD23_REQ_DONE || q->state == FD23_REQ_FAILED; };
    if (timeout_s < 0)
        s->cv_done.wait(lk, done);
    else
        s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
    if (state) *state = q->state;
    return FD23_OK;
}

extern "C" size_t fd23_server_result_bytes(fd23_server *s, long long id) {
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return 0;
    const Request &q = *it->second;
    return (size_t)q.r.nshots * q.ncomp * q.nframes * q.r.nrec * q.csize;
}

extern "C" int fd23_server_result(fd23_server *s, long long id, void *out, size_t bytes, double *latency_s,
                                  double *queue_s) {
    if (!s) {
        fd23_set_error_("null server");
        return FD23_EINVAL;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) {
        fd23_set_error_("unknown request id");
        return FD23_EINVAL;
    }
    auto q = it->second;
    if (q->state != FD23_REQ_DONE && q->state != FD23_REQ_FAILED) {
        fd23_set_error_("request not finished");
        return FD23_EINVAL;
    }
    if (latency_s) *latency_s = seconds(q->t_submit, q->t_done);
    if (queue_s) *queue_s = q->state == FD23_REQ_FAILED && q->t_admit == Clock::time_point{}
                                ? seconds(q->t_submit, q->t_done)
                                : seconds(q->t_submit, q->t_admit);
    if (q->state == FD23_REQ_FAILED) {
        fd23_set_error_(q->msg.c_str());
        const int e = q->err ? q->err : FD23_EHIP;
        s->all.erase(it);
        return e;
    }
    if (bytes != q->seis.size() || (bytes && !out)) {
        fd23_set_error_("result buffer size does not match fd23_server_result_bytes");
        return FD23_EINVAL;
    }
    if (bytes) std::memcpy(out, q->seis.data(), bytes);
    s->all.erase(it);
    return FD23_OK;
}

extern "C" int fd23_server_stats_get(fd23_server *s, fd23_server_stats *out) {
    if (!s || !out) {
        fd23_set_error_("null argument");
        return FD23_EINVAL;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->queued = (long long)s->queue.size();
    long long running = 0;
    for (auto &kv : s->all)
        if (kv.second->state == FD23_REQ_RUNNING) running++;
    out->running = running;
    out->bytes_in_use = s->used;
    return FD23_OK;
}
