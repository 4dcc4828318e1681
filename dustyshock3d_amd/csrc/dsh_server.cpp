// dsh_server.cpp - continuous batching of shock-tube requests on one GPU.
//
// The running set is one engine batch.  A scheduling round (pump):
//   1. admit queued requests in arrival order while the new batch and the old one
//      (both live during the rebuild) fit the HBM budget and, for a heavy request, fewer
//      than max_heavy heavy ones run; a request larger than the budget is rejected,
//      otherwise the queue waits at the first that cannot start;
//   2. if requests were admitted or cancelled, build a new batch: running tubes are
//      adopted device to device (state, sorted order, clock), new tubes get their
//      seeded set-up, finished tubes are dropped;
//   3. advance the batch steps_per_epoch steps (one graph launch per step);
//   4. retire finished tubes: copy their arrays to host memory.  Their slots stay in
//      the batch (the kernels skip a finished tube) until the next rebuild, so a tube
//      that finishes costs no rebuild of its own; when none is left running the batch
//      is released.
// The device behind this is a Backend: the GPU engine, or a host-only fake used to
// test the scheduling logic without a GPU.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dsh.h"
#include "dsh_engine.h"

namespace dsh {

int upload_tube_spec(Engine& eng, int tube, const dsh_tube_spec& spec);

namespace {

// Device bytes of one tube in a batch (the engine's allocation per particle, cell and
// averaging cell, including the second state buffer and the stored hit masks).
double tube_bytes(const dsh_tube_spec& spec) {
    int n[3];
    if (dsh_tube_count(&spec, n)) return 0.0;
    dsh_params p;
    if (dsh_tube_params(&spec, &p)) return 0.0;
    // smallest smoothing length of the set-up: the x <= 0 lattice
    const int nyr = std::max(1, (int)std::floor(spec.y_width / (2.0 * spec.dx_left) + 0.5));
    const int nzr = std::max(1, (int)std::floor(spec.z_width / (2.0 * spec.dx_left) + 0.5));
    const double hmin = 1.2 * std::cbrt(spec.dx_left * (spec.y_width / (2 * nyr)) * (spec.z_width / (2 * nzr)));
    TubeConst C{};
    build_grid(p, hmin, C);
    const double particles = (double)n[0] + (double)n[1];
    const double per_particle = 2.0 * 16 * 8 + 2.0 * 2 * 4 + 6 * 4 + 3 * 8;
    const double masks = std::ceil(n[0] / 64.0) * 48.0 * 64.0 * 12.0;   // hit masks + x-windows
    const double cells = 2.0 * C.ncell * (4 + 4 + 8 + 8 + 16);
    const double avg = 40.0 * C.navg;
    return particles * per_particle + masks + cells + avg + 4096.0;
}

// Estimated steps of a request: max_steps, or t_end over the CFL step of the set-up at rest
// (0.3 h / cs with h = 1.2 dx); the CFL step shrinks as the flow speeds up, so this is a lower
// bound that orders requests by work.
double est_steps(const dsh_tube_spec& spec) {
    if (spec.max_steps > 0) return (double)spec.max_steps;
    const double dt0 = spec.cfl * 1.2 * spec.dx_left;
    return dt0 > 0 && spec.t_end < 1e29 ? std::ceil(spec.t_end / dt0) : 1e30;
}

struct Request {
    uint64_t id = 0;
    dsh_tube_spec spec{};
    int state = DSH_REQ_QUEUED;
    double arrival = 0, admitted = -1, finished = -1, particle_steps = 0, bytes = 0;
    double work = 0;   // particles x estimated steps
    bool heavy = false;
    int n[2] = {0, 0};
    std::vector<double> f[2][16];
    std::vector<int> ind[2];
};

class Backend {
public:
    virtual ~Backend() = default;
    // New batch: slots [0, keep.size()) adopt the listed old slots, the rest start `add`.
    virtual int rebuild(const std::vector<int>& keep, const std::vector<const Request*>& add) = 0;
    virtual int advance(int steps) = 0;
    virtual int slot_state(int slot, int* done, int* steps, int* err) = 0;
    virtual int fetch(int slot, Request& r) = 0;
    virtual double bytes() const = 0;
    virtual std::string error() const = 0;
};

class GpuBackend : public Backend {
public:
    explicit GpuBackend(int variant) : variant_(variant) {}
    int rebuild(const std::vector<int>& keep, const std::vector<const Request*>& add) override {
        std::vector<dsh_params> p;
        for (int k : keep) p.push_back(params_[k]);
        for (const Request* r : add) {
            dsh_params q;
            if (dsh_tube_params(&r->spec, &q)) return fail(DSH_ERR_ARG, "bad tube spec");
            p.push_back(q);
        }
        if (p.empty()) {
            eng_.reset();
            params_.clear();
            bytes_ = 0;
            return DSH_OK;
        }
        std::unique_ptr<Engine> e(new Engine(p));
        int rc;
        for (size_t i = 0; i < keep.size(); ++i)
            if ((rc = e->adopt((int)i, *eng_, keep[i]))) return fail(rc, e->error());
        for (size_t i = 0; i < add.size(); ++i)
            if ((rc = upload_tube_spec(*e, (int)(keep.size() + i), add[i]->spec))) return fail(rc, e->error());
        e->set_pair_variant(variant_);
        if ((rc = e->prepare())) return fail(rc, e->error());
        eng_ = std::move(e);
        params_ = p;
        bytes_ = (double)eng_->device_bytes();
        return DSH_OK;
    }
    int advance(int steps) override {
        if (!eng_) return DSH_OK;
        int rc;
        if ((rc = eng_->launch_steps(steps))) return fail(rc, eng_->error());
        if ((rc = eng_->sync())) return fail(rc, eng_->error());
        return DSH_OK;
    }
    int slot_state(int slot, int* done, int* steps, int* err) override {
        const TubeDyn& D = eng_->dyn(slot);
        *done = D.done;
        *steps = D.steps;
        *err = D.err;
        return DSH_OK;
    }
    int fetch(int slot, Request& r) override {
        for (int s = 0; s < 2; ++s) {
            const size_t n = (size_t)r.n[s];
            for (auto& v : r.f[s]) v.assign(n, 0.0);
            r.ind[s].assign(n, 0);
            dsh_soa_view v{};
            double** dst[16] = {&v.x, &v.y, &v.z, &v.vx, &v.vy, &v.vz, &v.ax, &v.ay,
                                &v.az, &v.m, &v.rho, &v.p, &v.e, &v.h, &v.t_stop, &v.edot};
            for (int k = 0; k < 16; ++k) *dst[k] = r.f[s][k].data();
            v.ind = r.ind[s].data();
            const int rc = eng_->download(slot, s, &v);
            if (rc) return fail(rc, eng_->error());
        }
        return DSH_OK;
    }
    double bytes() const override { return bytes_; }
    std::string error() const override { return err_; }

private:
    int fail(int rc, const std::string& m) {
        err_ = m;
        return rc;
    }
    int variant_;
    std::unique_ptr<Engine> eng_;
    std::vector<dsh_params> params_;
    double bytes_ = 0;
    std::string err_;
};

// Host-only stand-in: a tube needs ceil(t_end / dt0) steps, dt0 the CFL step of rest.
class FakeBackend : public Backend {
public:
    explicit FakeBackend(int fail_at) : fail_at_(fail_at) {}
    int rebuild(const std::vector<int>& keep, const std::vector<const Request*>& add) override {
        ++rebuilds_;
        if (fail_at_ > 0 && rebuilds_ == fail_at_) {
            err_ = "fake device: allocation failed";
            return DSH_ERR_NOMEM;
        }
        std::vector<Slot> s;
        for (int k : keep) s.push_back(slots_[k]);
        for (const Request* r : add) {
            Slot x;
            x.need = (int)est_steps(r->spec);
            x.bytes = tube_bytes(r->spec);
            s.push_back(x);
        }
        slots_ = s;
        return DSH_OK;
    }
    int advance(int steps) override {
        for (auto& s : slots_) s.steps = std::min(s.need, s.steps + steps);
        return DSH_OK;
    }
    int slot_state(int slot, int* done, int* steps, int* err) override {
        *done = slots_[slot].steps >= slots_[slot].need;
        *steps = slots_[slot].steps;
        *err = 0;
        return DSH_OK;
    }
    int fetch(int, Request&) override { return DSH_OK; }
    double bytes() const override {
        double b = 0;
        for (auto& s : slots_) b += s.bytes;
        return b;
    }
    std::string error() const override { return err_; }

private:
    struct Slot {
        int need = 0, steps = 0;
        double bytes = 0;
    };
    std::vector<Slot> slots_;
    int fail_at_, rebuilds_ = 0;
    std::string err_;
};

double wall_clock(void*) {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

class Server {
public:
    Server(const dsh_server_opts& o, std::unique_ptr<Backend> be) : o_(o), be_(std::move(be)) {
        if (!o_.clock) o_.clock = wall_clock;
        if (o_.steps_per_epoch < 1) o_.steps_per_epoch = 1;
        if (o_.max_heavy < 1) o_.max_heavy = 1;
    }
    double now() { return o_.clock(o_.clock_user); }

    int submit(const dsh_tube_spec& spec, double arrival, uint64_t* id) {
        Request r;
        r.id = next_id_++;
        r.spec = spec;
        r.arrival = arrival;
        int n[3];
        if (dsh_tube_count(&spec, n)) return fail(DSH_ERR_ARG, "bad tube spec");
        r.n[0] = n[0];
        r.n[1] = n[1];
        r.bytes = tube_bytes(spec);
        r.work = (double)(n[0] + n[1]) * est_steps(spec);
        r.heavy = o_.heavy_work > 0 && r.work > o_.heavy_work;
        if (r.bytes > o_.hbm_budget || r.bytes > o_.hbm_capacity) r.state = DSH_REQ_REJECTED;
        reqs_[r.id] = std::move(r);
        if (reqs_[next_id_ - 1].state == DSH_REQ_QUEUED) queue_.push_back(next_id_ - 1);
        *id = next_id_ - 1;
        return DSH_OK;
    }

    int pump(int* running, int* queued) {
        // 1-2: admission and rebuild
        std::vector<int> keep;
        std::vector<uint64_t> keep_ids;
        double keep_bytes = 0;
        int heavy = 0;
        bool changed = false;
        for (size_t s = 0; s < slot_ids_.size(); ++s) {
            Request& r = reqs_[slot_ids_[s]];
            if (r.state == DSH_REQ_RUNNING) {
                keep.push_back((int)s);
                keep_ids.push_back(r.id);
                keep_bytes += r.bytes;
                heavy += r.heavy ? 1 : 0;
            } else if (r.state == DSH_REQ_CANCELLED) {
                changed = true;   // still stepped by the device: drop it now
            }
        }
        std::vector<const Request*> add;
        std::vector<uint64_t> add_ids;
        double new_bytes = keep_bytes;
        const double old_bytes = be_->bytes();
        while (!queue_.empty()) {
            Request& r = reqs_[queue_.front()];
            if (r.state != DSH_REQ_QUEUED) {
                queue_.pop_front();
                continue;
            }
            if (new_bytes + r.bytes > o_.hbm_budget || new_bytes + r.bytes + old_bytes > o_.hbm_capacity) break;
            if (r.heavy && heavy >= o_.max_heavy) break;
            heavy += r.heavy ? 1 : 0;
            new_bytes += r.bytes;
            add.push_back(&r);
            add_ids.push_back(r.id);
            queue_.pop_front();
        }
        if (!add.empty()) changed = true;
        if (changed) {
            const double t0 = wall_clock(nullptr);
            const int rc = be_->rebuild(keep, add);
            rebuild_s_ += wall_clock(nullptr) - t0;
            ++rebuilds_;
            if (rc != DSH_OK) {
                // the old batch keeps running; the new requests go back to the queue head
                for (auto it = add_ids.rbegin(); it != add_ids.rend(); ++it) queue_.push_front(*it);
                err_ = be_->error();
                if (rc != DSH_ERR_NOMEM) return fail(rc, err_);
            } else {
                slot_ids_ = keep_ids;
                const double t = now();
                for (uint64_t id : add_ids) {
                    reqs_[id].state = DSH_REQ_RUNNING;
                    reqs_[id].admitted = t;
                    slot_ids_.push_back(id);
                }
            }
        }
        // 3: advance
        if (!slot_ids_.empty()) {
            const double t0 = wall_clock(nullptr);
            const int rc = be_->advance(o_.steps_per_epoch);
            advance_s_ += wall_clock(nullptr) - t0;
            ++epochs_;
            if (rc) return fail(rc, be_->error());
        }
        // 4: retire
        const double t = now();
        for (size_t s = 0; s < slot_ids_.size(); ++s) {
            Request& r = reqs_[slot_ids_[s]];
            if (r.state != DSH_REQ_RUNNING) continue;
            int done = 0, steps = 0, err = 0;
            be_->slot_state((int)s, &done, &steps, &err);
            if (err) {
                r.state = DSH_REQ_FAILED;
                r.finished = t;
            } else if (done) {
                const int rc = be_->fetch((int)s, r);
                if (rc) return fail(rc, be_->error());
                r.state = DSH_REQ_DONE;
                r.finished = t;
                r.particle_steps = (double)(r.n[0] + r.n[1]) * steps;
            }
        }
        int nr = 0;
        for (uint64_t id : slot_ids_) nr += reqs_[id].state == DSH_REQ_RUNNING ? 1 : 0;
        if (nr == 0 && !slot_ids_.empty()) {
            // nothing left running: release the batch's device memory
            const double t0 = wall_clock(nullptr);
            const int rc = be_->rebuild({}, {});
            rebuild_s_ += wall_clock(nullptr) - t0;
            ++rebuilds_;
            if (rc) return fail(rc, be_->error());
            slot_ids_.clear();
        }
        if (running) *running = nr;
        if (queued) {
            int nq = 0;
            for (uint64_t id : queue_) nq += reqs_[id].state == DSH_REQ_QUEUED ? 1 : 0;
            *queued = nq;
        }
        return DSH_OK;
    }

    Request* find(uint64_t id) {
        auto it = reqs_.find(id);
        return it == reqs_.end() ? nullptr : &it->second;
    }
    int cancel(uint64_t id) {
        Request* r = find(id);
        if (!r) return DSH_ERR_NOTFOUND;
        if (r->state == DSH_REQ_QUEUED || r->state == DSH_REQ_RUNNING) {
            r->state = DSH_REQ_CANCELLED;
            r->finished = now();
        }
        return DSH_OK;
    }
    double bytes_in_use() const { return be_->bytes(); }
    int rebuilds() const { return rebuilds_; }
    int epochs() const { return epochs_; }
    double rebuild_seconds() const { return rebuild_s_; }
    double advance_seconds() const { return advance_s_; }
    std::string err_;

private:
    int fail(int rc, const std::string& m) {
        err_ = m;
        return rc;
    }
    dsh_server_opts o_;
    std::unique_ptr<Backend> be_;
    std::map<uint64_t, Request> reqs_;
    std::deque<uint64_t> queue_;
    std::vector<uint64_t> slot_ids_;
    uint64_t next_id_ = 1;
    int rebuilds_ = 0, epochs_ = 0;
    double rebuild_s_ = 0, advance_s_ = 0;
};

}  // namespace dsh

struct dsh_server {
    std::unique_ptr<dsh::Server> sv;
};

extern "C" {

int dsh_server_footprint(const dsh_tube_spec* spec, double* bytes) {
    if (!spec || !bytes) return DSH_ERR_ARG;
    int n[3];
    if (dsh_tube_count(spec, n)) return DSH_ERR_ARG;
    *bytes = dsh::tube_bytes(*spec);
    return DSH_OK;
}

int dsh_server_work(const dsh_tube_spec* spec, double* work) {
    if (!spec || !work) return DSH_ERR_ARG;
    int n[3];
    if (dsh_tube_count(spec, n)) return DSH_ERR_ARG;
    *work = (double)(n[0] + n[1]) * dsh::est_steps(*spec);
    return DSH_OK;
}

int dsh_server_create(const dsh_server_opts* o, dsh_server** out) {
    if (!o || !out || o->hbm_budget <= 0) return DSH_ERR_ARG;
    std::unique_ptr<dsh::Backend> be;
    dsh_server_opts opts = *o;
    if (o->backend == DSH_BACKEND_GPU) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return DSH_ERR_HIP;
        if (opts.hbm_capacity <= 0) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) != hipSuccess) return DSH_ERR_HIP;
            opts.hbm_capacity = 0.9 * (double)tot;
        }
        be.reset(new dsh::GpuBackend(o->pair_variant));
    } else if (o->backend == DSH_BACKEND_FAKE) {
        if (opts.hbm_capacity <= 0) opts.hbm_capacity = 2.0 * opts.hbm_budget;
        be.reset(new dsh::FakeBackend(o->fake_fail_rebuild));
    } else {
        return DSH_ERR_ARG;
    }
    dsh_server* s = new dsh_server();
    s->sv.reset(new dsh::Server(opts, std::move(be)));
    *out = s;
    return DSH_OK;
}

int dsh_server_submit(dsh_server* s, const dsh_tube_spec* spec, double arrival, uint64_t* id) {
    if (!s || !spec || !id) return DSH_ERR_ARG;
    return s->sv->submit(*spec, arrival, id);
}

int dsh_server_pump(dsh_server* s, int* running, int* queued) {
    if (!s) return DSH_ERR_ARG;
    return s->sv->pump(running, queued);
}

int dsh_server_poll(dsh_server* s, uint64_t id, int* state, double* arrival, double* admitted, double* finished,
                    double* particle_steps) {
    if (!s) return DSH_ERR_ARG;
    dsh::Request* r = s->sv->find(id);
    if (!r) return DSH_ERR_NOTFOUND;
    if (state) *state = r->state;
    if (arrival) *arrival = r->arrival;
    if (admitted) *admitted = r->admitted;
    if (finished) *finished = r->finished;
    if (particle_steps) *particle_steps = r->particle_steps;
    return DSH_OK;
}

int dsh_server_result(dsh_server* s, uint64_t id, int species, const dsh_soa_view* v) {
    if (!s || !v || species < 0 || species > 1) return DSH_ERR_ARG;
    dsh::Request* r = s->sv->find(id);
    if (!r) return DSH_ERR_NOTFOUND;
    if (r->state != DSH_REQ_DONE || r->f[species][0].empty()) return DSH_ERR_STATE;
    double* dst[16] = {v->x, v->y, v->z, v->vx, v->vy, v->vz, v->ax, v->ay,
                       v->az, v->m, v->rho, v->p, v->e, v->h, v->t_stop, v->edot};
    const size_t n = (size_t)r->n[species];
    for (int k = 0; k < 16; ++k)
        if (dst[k]) std::memcpy(dst[k], r->f[species][k].data(), n * sizeof(double));
    if (v->ind) std::memcpy(v->ind, r->ind[species].data(), n * sizeof(int));
    return DSH_OK;
}

int dsh_server_cancel(dsh_server* s, uint64_t id) {
    if (!s) return DSH_ERR_ARG;
    return s->sv->cancel(id);
}

int dsh_server_stats(dsh_server* s, double* bytes_in_use, int* rebuilds, int* epochs) {
    if (!s) return DSH_ERR_ARG;
    if (bytes_in_use) *bytes_in_use = s->sv->bytes_in_use();
    if (rebuilds) *rebuilds = s->sv->rebuilds();
    if (epochs) *epochs = s->sv->epochs();
    return DSH_OK;
}

int dsh_server_times(dsh_server* s, double* rebuild_s, double* advance_s) {
    if (!s) return DSH_ERR_ARG;
    if (rebuild_s) *rebuild_s = s->sv->rebuild_seconds();
    if (advance_s) *advance_s = s->sv->advance_seconds();
    return DSH_OK;
}

const char* dsh_server_last_error(dsh_server* s) { return s ? s->sv->err_.c_str() : "null server"; }

void dsh_server_destroy(dsh_server* s) { delete s; }

}  // extern "C"
