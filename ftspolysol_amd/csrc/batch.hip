// batch.hip -- replica batches: many independent CL systems stepping together.
//
// A single 96^3 system cannot fill the chip during its propagator sweeps: one
// plane pass is 96 workgroups on 256 CUs and each bead step waits for the one
// before it.  A batch groups replicas by mesh, precision and grid spacing; the
// replicas of a group share a stream, and every bead-step launch of a group
// carries the jobs (sweep directions) of all its replicas at that bead index,
// read from a device-resident job table built once.  Replicas with shorter
// chains drop out of the later launches.  Groups run concurrently, each from
// its own host thread on its own stream.  Everything outside the sweeps (the
// smoothing FFTs, ions, drifts and the CL update) is each replica's own work, on
// side streams forked from the group stream.  A whole CL step of a group (its
// sweeps, every replica's work and the copies of the Q values and finiteness flags
// to pinned memory) is captured once as a hipGraph and replayed: every argument of
// its launches is a fixed device address except the noise key {seed, step}, which
// the kernels read from device memory that a captured copy refreshes from pinned
// host memory on every replay.  The host waits once per step.
#include <memory>
#include <thread>

#include "engine.h"

namespace {

enum { L_PLANE = 0, L_PENCIL = 1, L_LINE = 2 };

struct SweepLaunch {
    int kind, mode, axis;
    size_t off;
    int count;
};

// The bead-step launches of one species index over a group: INIT, then per bead
// index s the launches of that step; all jobs live in three device arrays.
struct SweepTables {
    int p = -1, sched = 0, steps = 0;
    std::vector<char> mask; // participating replicas of the group
    std::vector<SweepLaunch> launches;
    size_t init_launches = 0;
    long long jobs = 0;      // bead steps (jobs summed over the step launches)
    void* dplane = nullptr;
    void* dpencil = nullptr;
    void* dline = nullptr;
    // the step launches [init_launches, launches.size()) captured once as a graph:
    // their arguments are the fixed device tables above, so every replay is the same work
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~SweepTables()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void* q : {dplane, dpencil, dline})
            if (q) (void)hipFree(q);
    }
};

// One captured CL step of a group: valid while the participating replicas, the
// replicas that record, the learned kernel's ping-pong state and every replica's
// configuration are those of the capture.
struct StepGraph {
    std::vector<char> mask, rec;
    std::vector<int> lrk_par, lrk_prev;
    std::vector<unsigned long long> epoch;
    bool timing = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // host state of each replica after the captured step, restored after every replay
    std::vector<const double2*> psi_hat_of, dens_psi;
    std::vector<int> last_species, par_after, prev_after;
    fts_timing dt{}; // timing counters one step adds
    int nstamp = 0;
    ~StepGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

constexpr int MAX_STAMPS = 2 * 2 * FTS_MAX_SPECIES; // two density evaluations (PC) x species x {before, after}

struct Group {
    int m1 = 0, m2 = 0, m3 = 0, precision = 0;
    double dL = 0;
    hipStream_t stream = nullptr;
    std::vector<fts_engine*> reps;
    std::vector<int> ids; // batch index of each replica
    int sched = FTS_SCHED_CONV;
    bool plane_ok = true;
    int graph_level = 2;    // fts_batch_set_graph: 0 eager launches, 1 sweep graphs, 2 whole-step graphs
    bool capturing = false; // inside a whole-step capture: sweeps launch their kernels into it
    std::vector<std::unique_ptr<StepGraph>> step_graphs;
    double tune_ms[2] = {0.0, 0.0};
    std::vector<std::unique_ptr<SweepTables>> tables; // full sweeps, per species index and schedule
    double2* hQ = nullptr;                             // pinned [2][reps][Q_SLOTS]: Q, Q at (w_bar, psi_bar)
    int* hflag = nullptr;                              // pinned [reps]
    double2* hdrift = nullptr;                         // pinned [reps][FTS_MAX_TYPES + 1]: drift statistics of the step
    // The work of single replicas (smoothing FFTs, ions, contractions, updates)
    // runs on side streams forked from the group stream and joined back into it,
    // so that the small launches of several replicas share the chip; replica r
    // takes side[r % side.size()].  No side streams: everything on the group stream.
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> joined; // per side stream
    hipEvent_t forked = nullptr;
    // timing: device wall-clock stamps before and after each sweep span of a step (HIP
    // events recorded inside a captured graph cannot be timed), copied to pinned memory
    // with the step's other results and summed after its sync
    bool timing = false;
    unsigned long long* dstamps = nullptr;
    unsigned long long* hstamps = nullptr;
    int nstamp = 0;
    double clock_khz = 0.0;
    fts_timing tstat{};
    int rc = FTS_OK;
    std::string err;
};

template <typename T> struct JobArrays {
    std::vector<PlaneJob<T>> plane;
    std::vector<PencilJob<T>> pencil;
    std::vector<LineJob<T>> line;
};

// per-replica pointers of one species' sweep
template <typename T> struct RepView {
    fts_engine* e;
    const HostSpecies* sp;
    V2<T>*qf, *qb, *E, *tf, *tb;
    double *nF[2], *nB[2], *lcF, *lcB;
    size_t Ms;
    V2<T>* E_of(int s) const { return E + static_cast<size_t>(sp->cls_of_bead[static_cast<size_t>(s)]) * Ms; }
    KOp bond_op(int bond) const
    {
        const size_t ks = static_cast<size_t>(e->m1 + e->m2 + e->m3);
        return e->kop_sep(e->sep_bond.p + static_cast<size_t>(sp->bond_cls[static_cast<size_t>(bond)]) * ks, 1.0 / static_cast<double>(e->M));
    }
    KOp bond_unit(int bond) const { return e->kop_sep(bond_op(bond).sep, 1.0); } // the factors without 1/M

};

template <typename T> RepView<T> view(fts_engine* e, int p)
{
    using Cv = V2<T>;
    RepView<T> v;
    v.e = e;
    v.sp = &e->species[static_cast<size_t>(p)];
    v.qf = static_cast<Cv*>(e->qF);
    v.qb = static_cast<Cv*>(e->qB);
    v.E = static_cast<Cv*>(e->E);
    v.tf = static_cast<Cv*>(e->tmpF);
    v.tb = static_cast<Cv*>(e->tmpB);
    v.Ms = static_cast<size_t>(e->M);
    v.nF[0] = e->norms.p;
    v.nF[1] = e->norms.p + e->norm_stride;
    v.nB[0] = e->norms.p + 2 * e->norm_stride;
    v.nB[1] = e->norms.p + 3 * e->norm_stride;
    v.lcF = e->logc.p;
    v.lcB = e->logc.p + e->Nmax;
    return v;
}

// copies on the group's stream: a copy on the legacy stream from one group's
// thread while another group captures its sweeps would be refused
template <typename X> int upload(const std::vector<X>& v, void** dst, hipStream_t stream)
{
    if (v.empty()) return FTS_OK;
    HIP_TRY(hipMalloc(dst, sizeof(X) * v.size()));
    HIP_TRY(hipMemcpyAsync(*dst, v.data(), sizeof(X) * v.size(), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return FTS_OK;
}

// Job tables of species index p over the group for max_steps bead steps (all if
// max_steps < 0).  The bead-step chain of run_sweeps, with the jobs of every
// replica whose chain still has bead index s + 1 in launch s.  A replica whose
// chain ends before the group's longest takes the non-final pass form on its
// last step (the extra forward FFT writes only its scratch field).
template <typename T> int build_tables(Group& g, int p, int sched, int max_steps, const std::vector<char>& mask, SweepTables& tb)
{
    std::vector<RepView<T>> act;
    for (size_t r = 0; r < g.reps.size(); ++r) {
        fts_engine* e = g.reps[r];
        if (mask[r] && static_cast<int>(e->species.size()) > p && e->species[static_cast<size_t>(p)].N > 1) act.push_back(view<T>(e, p));
    }
    tb.mask = mask;
    std::stable_sort(act.begin(), act.end(), [](const RepView<T>& a, const RepView<T>& b) { return a.sp->N > b.sp->N; });
    tb.p = p;
    tb.sched = sched;
    tb.launches.clear();
    tb.jobs = 0;
    if (act.empty()) {
        tb.steps = 0;
        return FTS_OK;
    }
    const int maxN = act[0].sp->N;
    const int nsteps = max_steps < 0 ? maxN - 1 : std::min(maxN - 1, max_steps);
    tb.steps = nsteps;
    JobArrays<T> J;
    auto add = [&](int kind, int mode, int axis, size_t off, size_t end) {
        tb.launches.push_back(SweepLaunch{kind, mode, axis, off, static_cast<int>(end - off)});
    };
    if (sched == FTS_SCHED_PLANE || sched_is_conv(sched)) {
        size_t o = J.plane.size();
        for (const RepView<T>& r : act) {
            const int N = r.sp->N;
            J.plane.push_back(PlaneJob<T>{nullptr, r.tf, r.qf, r.E_of(0), nullptr, r.nF[0], r.lcF});
            if (!r.sp->mirror)
                J.plane.push_back(PlaneJob<T>{nullptr, r.tb, r.qb + static_cast<size_t>(N - 1) * r.Ms, r.E_of(N - 1), nullptr, r.nB[0], r.lcB + (N - 1)});
        }
        add(L_PLANE, PLANE_INIT, 0, o, J.plane.size());
        tb.init_launches = tb.launches.size();
    }
    if (sched_is_conv(sched)) {
        // bead step s: pencil q_s -> tmp, plane tmp -> q_{s+1}; a replica whose chain
        // has ended drops out of both launches
        for (int s = 0; s < nsteps; ++s) {
            const size_t op = J.pencil.size();
            for (const RepView<T>& r : act) {
                const int N = r.sp->N;
                if (N - 1 <= s) break;
                J.pencil.push_back(PencilJob<T>{r.tf, r.bond_op(s), r.qf + static_cast<size_t>(s) * r.Ms});
                if (!r.sp->mirror) J.pencil.push_back(PencilJob<T>{r.tb, r.bond_op(N - 2 - s), r.qb + static_cast<size_t>(N - 1 - s) * r.Ms});
            }
            add(L_PENCIL, conv_pencil_mode(sched), 0, op, J.pencil.size());
            const size_t ol = J.plane.size();
            for (const RepView<T>& r : act) {
                const int N = r.sp->N;
                if (N - 1 <= s) break;
                const int sf = s + 1, sb = N - 2 - s;
                J.plane.push_back(PlaneJob<T>{r.tf, nullptr, r.qf + static_cast<size_t>(sf) * r.Ms, r.E_of(sf), r.nF[s % 2], r.nF[(s + 1) % 2], r.lcF + sf, r.bond_op(s)});
                if (!r.sp->mirror)
                    J.plane.push_back(PlaneJob<T>{r.tb, nullptr, r.qb + static_cast<size_t>(sb) * r.Ms, r.E_of(sb), r.nB[s % 2], r.nB[(s + 1) % 2], r.lcB + sb, r.bond_op(sb)});
            }
            tb.jobs += static_cast<long long>(J.plane.size() - ol);
            add(L_PLANE, PLANE_CONV_E, 0, ol, J.plane.size());
        }
    } else if (sched == FTS_SCHED_PLANE) {
        for (int s = 0; s < nsteps; ++s) {
            const size_t op = J.pencil.size();
            for (const RepView<T>& r : act) {
                const int N = r.sp->N;
                if (N - 1 <= s) break;
                J.pencil.push_back(PencilJob<T>{r.tf, r.bond_op(s)});
                if (!r.sp->mirror) J.pencil.push_back(PencilJob<T>{r.tb, r.bond_op(N - 2 - s)});
            }
            add(L_PENCIL, PENCIL_FWD_MUL_INV, 0, op, J.pencil.size());
            const size_t ol = J.plane.size();
            for (const RepView<T>& r : act) {
                const int N = r.sp->N;
                if (N - 1 <= s) break;
                const int sf = s + 1, sb = N - 2 - s;
                J.plane.push_back(PlaneJob<T>{r.tf, r.tf, r.qf + static_cast<size_t>(sf) * r.Ms, r.E_of(sf), r.nF[s % 2], r.nF[(s + 1) % 2], r.lcF + sf});
                if (!r.sp->mirror)
                    J.plane.push_back(PlaneJob<T>{r.tb, r.tb, r.qb + static_cast<size_t>(sb) * r.Ms, r.E_of(sb), r.nB[s % 2], r.nB[(s + 1) % 2], r.lcB + sb});
            }
            tb.jobs += static_cast<long long>(J.plane.size() - ol);
            add(L_PLANE, s == maxN - 2 ? PLANE_LAST : PLANE_STEP, 0, ol, J.plane.size());
        }
    } else if (sched == FTS_SCHED_SEPARABLE) {
        size_t o = J.line.size();
        for (const RepView<T>& r : act) {
            const int N = r.sp->N;
            J.line.push_back(LineJob<T>{nullptr, r.qf, r.E_of(0), KOp{}, nullptr, r.nF[0], r.lcF});
            if (!r.sp->mirror)
                J.line.push_back(LineJob<T>{nullptr, r.qb + static_cast<size_t>(N - 1) * r.Ms, r.E_of(N - 1), KOp{}, nullptr, r.nB[0], r.lcB + (N - 1)});
        }
        add(L_LINE, LINE_E, 1, o, J.line.size());
        tb.init_launches = tb.launches.size();
        // bead step s: x pass q_s -> tmp, z pass in place, y pass tmp -> q_{s+1}; a
        // replica whose chain has ended drops out of all three
        for (int s = 0; s < nsteps; ++s) {
            size_t nact = 0;
            while (nact < act.size() && act[nact].sp->N - 1 > s) ++nact;
            const size_t ox = J.line.size();
            for (size_t k = 0; k < nact; ++k) {
                const RepView<T>& r = act[k];
                const int N = r.sp->N;
                J.line.push_back(LineJob<T>{r.tf, nullptr, nullptr, r.bond_op(s), nullptr, nullptr, nullptr, r.qf + static_cast<size_t>(s) * r.Ms});
                if (!r.sp->mirror)
                    J.line.push_back(LineJob<T>{r.tb, nullptr, nullptr, r.bond_op(N - 2 - s), nullptr, nullptr, nullptr, r.qb + static_cast<size_t>(N - 1 - s) * r.Ms});
            }
            add(L_LINE, LINE_CONV, 0, ox, J.line.size());
            const size_t oz = J.line.size();
            for (size_t k = 0; k < nact; ++k) {
                const RepView<T>& r = act[k];
                J.line.push_back(LineJob<T>{r.tf, nullptr, nullptr, r.bond_unit(s)});
                if (!r.sp->mirror) J.line.push_back(LineJob<T>{r.tb, nullptr, nullptr, r.bond_unit(r.sp->N - 2 - s)});
            }
            add(L_LINE, LINE_CONV, 2, oz, J.line.size());
            const size_t oy = J.line.size();
            for (size_t k = 0; k < nact; ++k) {
                const RepView<T>& r = act[k];
                const int N = r.sp->N;
                const int sf = s + 1, sb = N - 2 - s;
                J.line.push_back(LineJob<T>{r.tf, r.qf + static_cast<size_t>(sf) * r.Ms, r.E_of(sf), r.bond_unit(s), r.nF[s % 2], r.nF[(s + 1) % 2], r.lcF + sf});
                if (!r.sp->mirror)
                    J.line.push_back(LineJob<T>{r.tb, r.qb + static_cast<size_t>(sb) * r.Ms, r.E_of(sb), r.bond_unit(sb), r.nB[s % 2], r.nB[(s + 1) % 2], r.lcB + sb});
            }
            tb.jobs += static_cast<long long>(J.line.size() - oy);
            add(L_LINE, LINE_CONV_E, 1, oy, J.line.size());
        }
    } else {
        size_t o = J.line.size();
        for (const RepView<T>& r : act) {
            const int N = r.sp->N;
            J.line.push_back(LineJob<T>{r.tf, r.qf, r.E_of(0), KOp{}, nullptr, r.nF[0], r.lcF});
            if (!r.sp->mirror)
                J.line.push_back(LineJob<T>{r.tb, r.qb + static_cast<size_t>(N - 1) * r.Ms, r.E_of(N - 1), KOp{}, nullptr, r.nB[0], r.lcB + (N - 1)});
        }
        add(L_LINE, LINE_E_FWD, 2, o, J.line.size());
        // plain y-pass jobs of the replicas active at each step: the prefix of ybase
        std::vector<size_t> ybase_end;
        const size_t ybase = J.line.size();
        for (const RepView<T>& r : act) {
            J.line.push_back(LineJob<T>{r.tf});
            if (!r.sp->mirror) J.line.push_back(LineJob<T>{r.tb});
            ybase_end.push_back(J.line.size());
        }
        add(L_LINE, LINE_FWD, 1, ybase, J.line.size());
        tb.init_launches = tb.launches.size();
        for (int s = 0; s < nsteps; ++s) {
            size_t nact = 0;
            while (nact < act.size() && act[nact].sp->N - 1 > s) ++nact;
            const size_t yend = ybase_end[nact - 1];
            const size_t ox = J.line.size();
            for (size_t k = 0; k < nact; ++k) {
                const RepView<T>& r = act[k];
                J.line.push_back(LineJob<T>{r.tf, nullptr, nullptr, r.bond_op(s)});
                if (!r.sp->mirror) J.line.push_back(LineJob<T>{r.tb, nullptr, nullptr, r.bond_op(r.sp->N - 2 - s)});
            }
            add(L_LINE, LINE_FWD_MUL_INV, 0, ox, J.line.size());
            add(L_LINE, LINE_INV, 1, ybase, yend);
            const bool last = s == maxN - 2;
            const size_t oz = J.line.size();
            for (size_t k = 0; k < nact; ++k) {
                const RepView<T>& r = act[k];
                const int N = r.sp->N;
                const int sf = s + 1, sb = N - 2 - s;
                J.line.push_back(LineJob<T>{r.tf, r.qf + static_cast<size_t>(sf) * r.Ms, r.E_of(sf), KOp{}, r.nF[s % 2], r.nF[(s + 1) % 2], r.lcF + sf});
                if (!r.sp->mirror)
                    J.line.push_back(LineJob<T>{r.tb, r.qb + static_cast<size_t>(sb) * r.Ms, r.E_of(sb), KOp{}, r.nB[s % 2], r.nB[(s + 1) % 2], r.lcB + sb});
            }
            tb.jobs += static_cast<long long>(J.line.size() - oz);
            add(L_LINE, last ? LINE_INV_E : LINE_INV_E_FWD, 2, oz, J.line.size());
            if (!last) add(L_LINE, LINE_FWD, 1, ybase, yend);
        }
    }
    int rc;
    if ((rc = upload(J.plane, &tb.dplane, g.stream)) || (rc = upload(J.pencil, &tb.dpencil, g.stream)) ||
        (rc = upload(J.line, &tb.dline, g.stream)))
        return rc;
    return FTS_OK;
}

template <typename T> int run_tables(Group& g, const SweepTables& tb, size_t first, size_t last)
{
    fts_engine* e0 = g.reps[0];
    const FftGeom& geo = sizeof(T) == 8 ? e0->g64 : e0->g32;
    const bool d = sizeof(T) == 8;
    const void* tw[3] = {d ? static_cast<const void*>(e0->tw64_x.p) : static_cast<const void*>(e0->tw32_x.p),
                         d ? static_cast<const void*>(e0->tw64_yz.p) : static_cast<const void*>(e0->tw32_yz.p),
                         d ? static_cast<const void*>(e0->tw64_z.p) : static_cast<const void*>(e0->tw32_z.p)};
    for (size_t i = first; i < last; ++i) {
        const SweepLaunch& L = tb.launches[i];
        if (L.count <= 0) continue;
        switch (L.kind) {
        case L_PLANE: HIP_TRY(launch_plane_table<T>(geo, L.mode, static_cast<const PlaneJob<T>*>(tb.dplane) + L.off, L.count, g.stream)); break;
        case L_PENCIL: HIP_TRY(launch_pencil_table<T>(geo, L.mode, static_cast<const PencilJob<T>*>(tb.dpencil) + L.off, L.count, g.stream)); break;
        default: HIP_TRY(launch_line_table<T>(geo, L.axis, L.mode, tw[L.axis], static_cast<const LineJob<T>*>(tb.dline) + L.off, L.count, g.stream));
        }
    }
    return FTS_OK;
}

// device wall-clock stamp (constant-rate counter of hipDeviceAttributeWallClockRate kHz)
__global__ void k_stamp(unsigned long long* __restrict__ out)
{
    if (threadIdx.x == 0) out[0] = wall_clock64();
}

// sweep time of the stamps written by the step just synchronised
void resolve_stamps(Group& g)
{
    for (int k = 0; k + 1 < g.nstamp; k += 2)
        g.tstat.sweep_ms += static_cast<double>(g.hstamps[k + 1] - g.hstamps[k]) / g.clock_khz;
    g.nstamp = 0;
}

int stamp(Group& g)
{
    if (g.nstamp >= MAX_STAMPS) return fail(FTS_DEVICE_ERROR, "batch: too many timed sweep spans in one step");
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, g.stream, g.dstamps + g.nstamp);
    HIP_TRY(hipGetLastError());
    ++g.nstamp;
    return FTS_OK;
}

// the job tables of species p over the replicas in mask (built once per participant set)
template <typename T> int ensure_tables(Group& g, int p, const std::vector<char>& mask, SweepTables** out)
{
    SweepTables* tb = nullptr;
    for (auto& t : g.tables)
        if (t->p == p && t->sched == g.sched && t->mask == mask) tb = t.get();
    if (!tb) {
        // tables are cached per participant set; sets change only around restarts
        if (g.tables.size() >= 32) {
            HIP_TRY(hipStreamSynchronize(g.stream));
            g.tables.clear();
        }
        g.tables.emplace_back(new SweepTables());
        tb = g.tables.back().get();
        int rc = build_tables<T>(g, p, g.sched, -1, mask, *tb);
        if (rc) return rc;
    }
    *out = tb;
    return FTS_OK;
}

template <typename T> int group_sweeps(Group& g, int p, const std::vector<char>& mask)
{
    SweepTables* tb = nullptr;
    int rc = ensure_tables<T>(g, p, mask, &tb);
    if (rc) return rc;
    if (tb->steps == 0) return FTS_OK;
    if ((rc = run_tables<T>(g, *tb, 0, tb->init_launches))) return rc;
    if (g.timing && (rc = stamp(g))) return rc;
    if (g.graph_level >= 1 && !g.capturing) {
        if (!tb->exec) {
            HIP_TRY(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
            rc = run_tables<T>(g, *tb, tb->init_launches, tb->launches.size());
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(g.stream, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            HIP_TRY(ec);
            tb->graph = graph;
            HIP_TRY(hipGraphInstantiate(&tb->exec, tb->graph, nullptr, nullptr, 0));
        }
        HIP_TRY(hipGraphLaunch(tb->exec, g.stream));
    } else if ((rc = run_tables<T>(g, *tb, tb->init_launches, tb->launches.size()))) {
        return rc;
    }
    if (g.timing) {
        if ((rc = stamp(g))) return rc;
        g.tstat.sweeps += 1;
        g.tstat.bead_steps += tb->jobs;
        g.tstat.step_launches += tb->steps;
        g.tstat.algorithmic_bytes += static_cast<double>(tb->jobs) * 2.0 * sizeof(V2<T>) * static_cast<double>(g.reps[0]->M);
    }
    return FTS_OK;
}

// Runs fn(r) for every replica r in mask with the replica's engine on its side
// stream.  The side streams start after the work queued so far on the group
// stream, and the group stream waits for all of them afterwards, so the replicas'
// work is ordered as if it ran on the group stream.
template <typename F> int group_each(Group& g, const std::vector<char>& mask, F&& fn)
{
    int rc = FTS_OK;
    const size_t S = g.side.size();
    if (S == 0) {
        for (size_t r = 0; r < g.reps.size() && !rc; ++r)
            if (mask[r]) rc = fn(r);
        return rc;
    }
    HIP_TRY(hipEventRecord(g.forked, g.stream));
    std::vector<char> used(S, 0);
    for (size_t r = 0; r < g.reps.size() && !rc; ++r) {
        if (!mask[r]) continue;
        const size_t k = r % S;
        if (!used[k]) {
            HIP_TRY(hipStreamWaitEvent(g.side[k], g.forked, 0));
            used[k] = 1;
        }
        fts_engine* e = g.reps[r];
        e->stream = g.side[k];
        rc = fn(r);
        e->stream = g.stream;
    }
    for (size_t k = 0; k < S; ++k)
        if (used[k]) {
            HIP_TRY(hipEventRecord(g.joined[k], g.side[k]));
            HIP_TRY(hipStreamWaitEvent(g.stream, g.joined[k], 0));
        }
    return rc;
}

// compute_densities of the replicas in mask at (w, psi) or (w_bar, psi_bar)
// tail (optional): replica r's work after its densities, in the same fork of side
// streams (one join fewer per step)
template <typename T, typename Tail> int group_densities(Group& g, bool bar, const std::vector<char>& mask, Tail&& tail)
{
    int rc;
    size_t max_species = 0;
    for (size_t r = 0; r < g.reps.size(); ++r)
        if (mask[r]) max_species = std::max(max_species, g.reps[r]->species.size());
    // replica r's own work before the sweeps of species p (p = 0: densities_begin
    // first) and after them (the last species: densities_end too)
    auto has = [&](size_t r, int p) { return static_cast<int>(g.reps[r]->species.size()) > p; };
    rc = group_each(g, mask, [&](size_t r) {
        fts_engine* e = g.reps[r];
        int c = e->densities_begin(bar ? e->omega_bar.p : e->omega.p, bar ? e->psi_bar.p : e->psi.p);
        if (!c && has(r, 0)) c = e->solve_begin<T>(0);
        return c;
    });
    if (rc) return rc;
    for (int p = 0; p < static_cast<int>(max_species); ++p) {
        if ((rc = group_sweeps<T>(g, p, mask))) return rc;
        const bool last = p + 1 == static_cast<int>(max_species);
        rc = group_each(g, mask, [&](size_t r) {
            fts_engine* e = g.reps[r];
            int c = has(r, p) ? e->solve_end<T>(p, true) : FTS_OK;
            if (!c && has(r, p + 1)) c = e->solve_begin<T>(p + 1);
            if (!c && last) c = e->densities_end();
            if (!c && last) c = tail(r);
            return c;
        });
        if (rc) return rc;
    }
    if (max_species == 0)
        return group_each(g, mask, [&](size_t r) {
            int c = g.reps[r]->densities_end();
            return c ? c : tail(r);
        });
    return FTS_OK;
}
template <typename T> int group_densities(Group& g, bool bar, const std::vector<char>& mask)
{
    return group_densities<T>(g, bar, mask, [](size_t) { return FTS_OK; });
}

// queue the copy of the replicas' Q into slot k of g.hQ (read after a sync)
int group_queue_Q(Group& g, int k, const std::vector<char>& mask)
{
    int rc;
    for (size_t r = 0; r < g.reps.size(); ++r)
        if (mask[r] && (rc = g.reps[r]->q_queue(g.hQ + (static_cast<size_t>(k) * g.reps.size() + r) * Q_SLOTS))) return rc;
    return FTS_OK;
}

bool group_Q_finite(const Group& g, int k, size_t r)
{
    return g.reps[r]->partition_finite(g.hQ + (static_cast<size_t>(k) * g.reps.size() + r) * Q_SLOTS);
}

// One CL step of the replicas in mask; rec[r] non-null where replica r records.
// The whole step is queued before the host waits once: Q (and Q at the predicted
// fields) and the finiteness of the updated fields are checked after that sync, in
// this order, so a replica with a non-finite Q has run the rest of its step on
// non-finite values, which the caller then discards.  The host therefore queues
// the update work while the device runs the sweeps.
template <typename T> int group_step_queue(Group& g, int scheme, const std::vector<fts_observables*>& rec, const std::vector<char>& mask)
{
    int rc;
    // the engines of a batch defer measure's host part (measure_finish) to after the sync
    auto update = [&](size_t r) {
        fts_engine* e = g.reps[r];
        int c;
        if (scheme == FTS_SCHEME_PC_REFERENCE) c = e->pc_correct(rec[r] != nullptr, rec[r]);
        else if (scheme_two_stage(scheme)) c = e->sipc_correct(rec[r] != nullptr, rec[r]);
        else c = e->em_update(rec[r] != nullptr, rec[r], e->si_scheme());
        if (c) return c;
        ++e->step_counter;
        if (e->drift_stats)
            HIP_TRY(hipMemcpyAsync(g.hdrift + r * (FTS_MAX_TYPES + 1), e->scal.p + SC_DRIFT, sizeof(double2) * (static_cast<size_t>(e->n_omega) + 1),
                                   hipMemcpyDeviceToHost, e->stream));
        return e->finite_queue(g.hflag + r);
    };
    if (scheme_two_stage(scheme)) {
        const bool pc = scheme == FTS_SCHEME_PC_REFERENCE;
        if ((rc = group_densities<T>(g, false, mask)) || (rc = group_queue_Q(g, 0, mask))) return rc;
        if ((rc = group_each(g, mask, [&](size_t r) { return pc ? g.reps[r]->pc_predict() : g.reps[r]->sipc_predict(); }))) return rc;
        if ((rc = group_densities<T>(g, true, mask)) || (rc = group_queue_Q(g, 1, mask))) return rc;
        if ((rc = group_each(g, mask, update))) return rc;
    } else {
        // one fork for each replica's densities, Q copy and update
        rc = group_densities<T>(g, false, mask, [&](size_t r) {
            int c = g.reps[r]->q_queue(g.hQ + r * Q_SLOTS);
            return c ? c : update(r);
        });
        if (rc) return rc;
    }
    if (g.timing && g.nstamp > 0)
        HIP_TRY(hipMemcpyAsync(g.hstamps, g.dstamps, sizeof(unsigned long long) * static_cast<size_t>(g.nstamp), hipMemcpyDeviceToHost, g.stream));
    return FTS_OK;
}

// After the step's sync: a replica whose step fails leaves it with its reason in why[r]
// (the caller restarts it or gives up); records finish on the host.
int group_step_check(Group& g, int scheme, const std::vector<fts_observables*>& rec, const std::vector<char>& mask,
                     std::vector<std::string>& why)
{
    int rc;
    const bool two_stage = scheme_two_stage(scheme); // Q at the predicted fields too
    if (g.timing) resolve_stamps(g);
    for (size_t r = 0; r < g.reps.size(); ++r) {
        if (!mask[r]) continue;
        const std::string at = " at step " + std::to_string(g.reps[r]->step_counter);
        const bool q_ok = group_Q_finite(g, 0, r) && (!two_stage || group_Q_finite(g, 1, r));
        if (!q_ok) --g.reps[r]->step_counter; // a step that fails at Q is not counted (as fts_step)
        if (!group_Q_finite(g, 0, r)) why[r] = "non-finite Q" + at;
        else if (two_stage && !group_Q_finite(g, 1, r)) why[r] = "non-finite Qbar" + at;
        else if (!g.hflag[r]) why[r] = "updated fields have non-finite values" + at;
        else if (rec[r] && (rc = g.reps[r]->measure_finish(rec[r]))) return rc;
    }
    return FTS_OK;
}

// Whether this step may run as a captured graph: level 2, and no replica with the
// per-step host work a replay would skip (drift statistics, the NaN test hook)
bool step_graph_ok(const Group& g, const std::vector<char>& mask)
{
    if (g.graph_level < 2) return false;
    for (size_t r = 0; r < g.reps.size(); ++r)
        if (mask[r] && (g.reps[r]->drift_stats || g.reps[r]->inject_step >= 0)) return false;
    return true;
}

StepGraph* find_step_graph(Group& g, const std::vector<char>& mask, const std::vector<fts_observables*>& rec)
{
    for (auto& sg : g.step_graphs) {
        if (sg->mask != mask || sg->timing != g.timing) continue;
        bool same = true;
        for (size_t r = 0; r < g.reps.size() && same; ++r) {
            const fts_engine* e = g.reps[r];
            same = sg->rec[r] == (rec[r] != nullptr ? 1 : 0) && sg->epoch[r] == e->config_epoch &&
                   (!mask[r] || (sg->lrk_par[r] == e->lrk_par && sg->lrk_prev[r] == (e->lrk_have_prev ? 1 : 0)));
        }
        if (same) return sg.get();
    }
    return nullptr;
}

// One CL step of the group as a replayed graph (captured at its first use), or eagerly
template <typename T> int group_step(Group& g, int scheme, const std::vector<fts_observables*>& rec, std::vector<char> mask,
                                     std::vector<std::string>& why)
{
    int rc;
    g.nstamp = 0;
    if (!step_graph_ok(g, mask)) {
        if ((rc = group_step_queue<T>(g, scheme, rec, mask))) return rc;
        HIP_TRY(hipStreamSynchronize(g.stream));
        return group_step_check(g, scheme, rec, mask, why);
    }
    const size_t n = g.reps.size();
    StepGraph* sg = find_step_graph(g, mask, rec);
    if (!sg) {
        // the sweeps' job tables exist before the capture (building them uploads and waits)
        size_t max_species = 0;
        for (size_t r = 0; r < n; ++r)
            if (mask[r]) max_species = std::max(max_species, g.reps[r]->species.size());
        for (size_t p = 0; p < max_species; ++p) {
            SweepTables* tb = nullptr;
            if ((rc = ensure_tables<T>(g, static_cast<int>(p), mask, &tb))) return rc;
        }
        if (g.step_graphs.size() >= 16) {
            HIP_TRY(hipStreamSynchronize(g.stream));
            g.step_graphs.clear();
        }
        auto fresh = std::make_unique<StepGraph>();
        fresh->mask = mask;
        fresh->timing = g.timing;
        for (size_t r = 0; r < n; ++r) {
            const fts_engine* e = g.reps[r];
            fresh->rec.push_back(rec[r] != nullptr ? 1 : 0);
            fresh->epoch.push_back(e->config_epoch);
            fresh->lrk_par.push_back(e->lrk_par);
            fresh->lrk_prev.push_back(e->lrk_have_prev ? 1 : 0);
        }
        const fts_timing before = g.tstat;
        HIP_TRY(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
        g.capturing = true;
        rc = group_step_queue<T>(g, scheme, rec, mask);
        g.capturing = false;
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(g.stream, &graph);
        fresh->graph = graph;
        if (rc) return rc;
        HIP_TRY(ec);
        HIP_TRY(hipGraphInstantiate(&fresh->exec, fresh->graph, nullptr, nullptr, 0));
        // the capture ran the step's host code: the replicas' host state is that after the step
        for (size_t r = 0; r < n; ++r) {
            const fts_engine* e = g.reps[r];
            fresh->psi_hat_of.push_back(e->psi_hat_of);
            fresh->dens_psi.push_back(e->dens_psi);
            fresh->last_species.push_back(e->last_species);
            fresh->par_after.push_back(e->lrk_par);
            fresh->prev_after.push_back(e->lrk_have_prev ? 1 : 0);
        }
        fresh->dt.sweeps = g.tstat.sweeps - before.sweeps;
        fresh->dt.bead_steps = g.tstat.bead_steps - before.bead_steps;
        fresh->dt.step_launches = g.tstat.step_launches - before.step_launches;
        fresh->dt.algorithmic_bytes = g.tstat.algorithmic_bytes - before.algorithmic_bytes;
        fresh->nstamp = g.nstamp;
        g.step_graphs.push_back(std::move(fresh));
        HIP_TRY(hipGraphLaunch(g.step_graphs.back()->exec, g.stream));
    } else {
        // replay: the step's noise keys and record headers, then the host state the
        // captured step leaves behind
        for (size_t r = 0; r < n; ++r) {
            if (!mask[r]) continue;
            fts_engine* e = g.reps[r];
            e->hkey[0] = e->seed;
            e->hkey[1] = static_cast<unsigned long long>(e->step_counter);
            if (rec[r]) e->measure_header(rec[r]);
        }
        HIP_TRY(hipGraphLaunch(sg->exec, g.stream));
        for (size_t r = 0; r < n; ++r) {
            if (!mask[r]) continue;
            fts_engine* e = g.reps[r];
            ++e->step_counter;
            e->psi_hat_of = sg->psi_hat_of[r];
            e->dens_psi = sg->dens_psi[r];
            e->last_species = sg->last_species[r];
            e->lrk_par = sg->par_after[r];
            e->lrk_have_prev = sg->prev_after[r] != 0;
        }
        if (g.timing) {
            g.tstat.sweeps += sg->dt.sweeps;
            g.tstat.bead_steps += sg->dt.bead_steps;
            g.tstat.step_launches += sg->dt.step_launches;
            g.tstat.algorithmic_bytes += sg->dt.algorithmic_bytes;
        }
        g.nstamp = sg->nstamp;
    }
    HIP_TRY(hipStreamSynchronize(g.stream));
    return group_step_check(g, scheme, rec, mask, why);
}

// Schedule of a group's sweeps: conv and separable timed over 8 bead steps of species
// 0 with all replicas in each launch, best of three; tune_ms = {conv, separable}
// (plane and line only when asked for, as for single engines).
template <typename T> int group_autotune(Group& g)
{
    if (!g.plane_ok) {
        g.sched = FTS_SCHED_SEPARABLE;
        return FTS_OK;
    }
    SweepTables t[2];
    const std::vector<char> all(g.reps.size(), 1);
    int rc = FTS_OK;
    const int cand[2] = {FTS_SCHED_CONV, FTS_SCHED_SEPARABLE};
    for (int k = 0; k < 2 && rc == FTS_OK; ++k) rc = build_tables<T>(g, 0, cand[k], 8, all, t[k]);
    if (rc) return rc;
    if (t[0].steps == 0) return FTS_OK;
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    double best[2] = {1e30, 1e30};
    for (int rep = 0; rep < 3 && rc == FTS_OK; ++rep)
        for (int k = 0; k < 2 && rc == FTS_OK; ++k) {
            if (hipEventRecord(a, g.stream) != hipSuccess) rc = fail(FTS_DEVICE_ERROR, "hipEventRecord");
            if (!rc) rc = run_tables<T>(g, t[k], 0, t[k].launches.size());
            if (!rc && (hipEventRecord(b, g.stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess)) rc = fail(FTS_DEVICE_ERROR, "hipEvent");
            float ms = 0.0f;
            if (!rc && hipEventElapsedTime(&ms, a, b) == hipSuccess) best[k] = std::min(best[k], static_cast<double>(ms));
        }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    int pick = 0;
    for (int k = 0; k < 2; ++k) {
        g.tune_ms[k] = best[k];
        if (best[k] < best[pick]) pick = k;
    }
    g.sched = cand[pick];
    return FTS_OK;
}

// replaces the side streams of g with n new ones (n = 0: none)
int set_side_streams(Group& g, int n)
{
    if (g.stream) HIP_TRY(hipStreamSynchronize(g.stream));
    for (hipStream_t s : g.side) (void)hipStreamDestroy(s);
    for (hipEvent_t e : g.joined) (void)hipEventDestroy(e);
    g.side.clear();
    g.joined.clear();
    for (int k = 0; k < n; ++k) {
        hipStream_t s = nullptr;
        hipEvent_t e = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        g.side.push_back(s);
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        g.joined.push_back(e);
    }
    if (n > 0 && !g.forked) HIP_TRY(hipEventCreateWithFlags(&g.forked, hipEventDisableTiming));
    return FTS_OK;
}

} // namespace

struct fts_batch {
    std::vector<fts_engine*> reps;
    std::vector<int> group_of;
    std::vector<std::unique_ptr<Group>> groups;
    int scheme = 0;
    ~fts_batch()
    {
        for (auto& g : groups)
            if (g->stream) (void)hipStreamSynchronize(g->stream);
        for (fts_engine* e : reps) delete e; // engines do not own the group streams
        for (auto& g : groups) {
            g->step_graphs.clear();
            g->tables.clear();
            if (g->dstamps) (void)hipFree(g->dstamps);
            if (g->hstamps) (void)hipHostFree(g->hstamps);
            if (g->hQ) (void)hipHostFree(g->hQ);
            if (g->hdrift) (void)hipHostFree(g->hdrift);
            if (g->hflag) (void)hipHostFree(g->hflag);
            set_side_streams(*g, 0);
            if (g->forked) (void)hipEventDestroy(g->forked);
            if (g->stream) (void)hipStreamDestroy(g->stream);
        }
    }
};

extern "C" {

int fts_batch_create(const fts_system* systems, int n, fts_batch** out)
{
    if (!systems || !out || n <= 0) return fail(FTS_INVALID, "batch: null argument or no replicas");
    *out = nullptr;
    for (int r = 1; r < n; ++r)
        if (systems[r].scheme != systems[0].scheme) return fail(FTS_INVALID, "batch: every replica must use the same scheme");
    std::unique_ptr<fts_batch> b(new fts_batch());
    b->scheme = systems[0].scheme;
    for (int r = 0; r < n; ++r) {
        const fts_system& s = systems[r];
        Group* g = nullptr;
        for (auto& c : b->groups)
            if (c->m1 == s.m1 && c->m2 == s.m2 && c->m3 == s.m3 && c->precision == s.precision && c->dL == s.dL) g = c.get();
        if (!g) {
            b->groups.emplace_back(new Group());
            g = b->groups.back().get();
            g->m1 = s.m1;
            g->m2 = s.m2;
            g->m3 = s.m3;
            g->precision = s.precision;
            g->dL = s.dL;
            // non-blocking: the group captures its sweeps on this stream, and a blocking
            // stream under capture makes every legacy-stream call of the process fail
            HIP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
        }
        fts_engine* e = nullptr;
        int rc = engine_create(&s, g->stream, false, &e);
        if (rc) return fail(rc, "batch replica " + std::to_string(r) + ": " + g_error);
        b->reps.push_back(e);
        b->group_of.push_back(static_cast<int>(std::find_if(b->groups.begin(), b->groups.end(), [&](auto& c) { return c.get() == g; }) - b->groups.begin()));
        g->reps.push_back(e);
        g->ids.push_back(r);
    }
    for (auto& gp : b->groups) {
        Group& g = *gp;
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&g.hQ), sizeof(double2) * Q_SLOTS * 2 * g.reps.size()));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&g.hflag), sizeof(int) * g.reps.size()));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&g.hdrift), sizeof(double2) * (FTS_MAX_TYPES + 1) * g.reps.size()));
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&g.dstamps), sizeof(unsigned long long) * MAX_STAMPS));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&g.hstamps), sizeof(unsigned long long) * MAX_STAMPS));
        {
            int dev = 0, khz = 0;
            HIP_TRY(hipGetDevice(&dev));
            HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
            if (khz <= 0) return fail(FTS_DEVICE_ERROR, "batch: no device wall-clock rate");
            g.clock_khz = static_cast<double>(khz);
        }
        g.plane_ok = true;
        for (fts_engine* e : g.reps) g.plane_ok = g.plane_ok && e->plane_ok;
        const int want = systems[g.ids[0]].fft_schedule;
        int rc = FTS_OK;
        if (want == FTS_SCHED_AUTO) rc = g.precision == FTS_PREC_FP32 ? group_autotune<float>(g) : group_autotune<double>(g);
        else if ((want == FTS_SCHED_PLANE || sched_is_conv(want)) && !g.plane_ok) return fail(FTS_INVALID, "batch: the plane schedules do not fit this mesh");
        else g.sched = want;
        if (rc) return rc;
        for (fts_engine* e : g.reps) {
            e->sched = g.sched;
            for (int k = 0; k < 2; ++k) e->tune_ms[k] = g.tune_ms[k];
            e->defer_measure = true;
        }
        if ((rc = set_side_streams(g, std::min<int>(static_cast<int>(g.reps.size()), FTS_BATCH_DEFAULT_STREAMS)))) return rc;
    }
    *out = b.release();
    return FTS_OK;
}

void fts_batch_destroy(fts_batch* b) { delete b; }

int fts_batch_get_info(fts_batch* b, fts_batch_info* o)
{
    if (!b || !o) return fail(FTS_INVALID, "null argument");
    std::memset(o, 0, sizeof(*o));
    o->n_replicas = static_cast<int>(b->reps.size());
    o->n_groups = static_cast<int>(b->groups.size());
    for (size_t k = 0; k < b->groups.size() && k < FTS_MAX_BATCH_GROUPS; ++k) {
        const Group& g = *b->groups[k];
        o->group_size[k] = static_cast<int>(g.reps.size());
        o->group_mesh[k][0] = g.m1;
        o->group_mesh[k][1] = g.m2;
        o->group_mesh[k][2] = g.m3;
        o->group_precision[k] = g.precision;
        o->group_schedule[k] = g.sched;
        for (int j = 0; j < 2; ++j) o->group_tune_ms[k][j] = g.tune_ms[j];
    }
    return FTS_OK;
}

int fts_batch_replica(fts_batch* b, int r, fts_engine** engine, int* group)
{
    if (!b || r < 0 || r >= static_cast<int>(b->reps.size())) return fail(FTS_INVALID, "batch: replica index");
    if (engine) *engine = b->reps[static_cast<size_t>(r)];
    if (group) *group = b->group_of[static_cast<size_t>(r)];
    return FTS_OK;
}

int fts_batch_step(fts_batch* b, int n, int record_every, fts_observables* records, int capacity, int* n_records)
{
    if (!b || n < 0) return fail(FTS_INVALID, "batch: bad argument");
    const size_t R = b->reps.size();
    if (n_records)
        for (size_t r = 0; r < R; ++r) n_records[r] = 0;
    auto run_group = [&](Group& g) {
        g.rc = FTS_OK;
        const size_t n_rep = g.reps.size();
        std::vector<fts_observables*> rec(n_rep, nullptr);
        std::vector<long long> target(n_rep);
        for (size_t r = 0; r < n_rep; ++r) target[r] = g.reps[r]->step_counter + n;
        std::vector<char> mask(n_rep);
        std::vector<std::string> why(n_rep);
        std::vector<long long> label(n_rep);
        while (g.rc == FTS_OK) {
            bool any = false;
            for (size_t r = 0; r < n_rep; ++r) {
                mask[r] = g.reps[r]->step_counter < target[r];
                any = any || mask[r];
            }
            if (!any) break;
            for (size_t r = 0; r < n_rep && g.rc == FTS_OK; ++r) {
                rec[r] = nullptr;
                why[r].clear();
                if (!mask[r]) continue;
                g.rc = g.reps[r]->maybe_checkpoint();
                const int id = g.ids[r];
                const long long step = g.reps[r]->step_counter + 1; // 1-based label, as fts_step
                const int have = n_records ? n_records[id] : 0;
                const bool want = record_every > 0 && step % record_every == 0 && records && n_records && have < capacity;
                rec[r] = want ? records + static_cast<size_t>(id) * capacity + have : nullptr;
                label[r] = step;
            }
            if (g.rc) break;
            g.rc = g.precision == FTS_PREC_FP32 ? group_step<float>(g, b->scheme, rec, mask, why) : group_step<double>(g, b->scheme, rec, mask, why);
            if (g.rc) break;
            for (size_t r = 0; r < n_rep; ++r) {
                const int id = g.ids[r];
                fts_engine* e = g.reps[r];
                if (rec[r]) {
                    rec[r]->step = label[r];
                    ++n_records[id];
                }
                if (why[r].empty()) {
                    // drift statistics of an accepted step: [step, modes..., psi], as fts_step
                    if (mask[r] && e->drift_stats) {
                        e->drift_rows.push_back(static_cast<double>(label[r]));
                        for (int k = 0; k <= e->n_omega; ++k) e->drift_rows.push_back(g.hdrift[r * (FTS_MAX_TYPES + 1) + static_cast<size_t>(k)].x);
                    }
                    continue;
                }
                // non-finite: restart this replica from its checkpoint with new noise, or give up
                g_error = why[r];
                int rr = e->restart_from_checkpoint(label[r]);
                if (rr) {
                    g.rc = fail(rr, "replica " + std::to_string(id) + ": " + g_error);
                    break;
                }
                if (n_records)
                    while (n_records[id] > 0 && records[static_cast<size_t>(id) * capacity + n_records[id] - 1].step > e->ckpt_step) --n_records[id];
                const size_t w = static_cast<size_t>(e->n_omega) + 2;
                while (!e->drift_rows.empty() && e->drift_rows[e->drift_rows.size() - w] > static_cast<double>(e->ckpt_step))
                    e->drift_rows.resize(e->drift_rows.size() - w);
            }
        }
        if (g.rc) g.err = g_error;
    };
    if (b->groups.size() == 1) {
        run_group(*b->groups[0]);
    } else {
        std::vector<std::thread> th;
        for (auto& g : b->groups) th.emplace_back(run_group, std::ref(*g));
        for (auto& t : th) t.join();
    }
    for (auto& g : b->groups)
        if (g->rc) return fail(g->rc, g->err);
    return FTS_OK;
}

int fts_batch_set_streams(fts_batch* b, int n)
{
    if (!b || n < 0 || n > FTS_BATCH_MAX_STREAMS) return fail(FTS_INVALID, "batch: side stream count out of range");
    for (auto& g : b->groups) {
        int rc = set_side_streams(*g, std::min<int>(n, static_cast<int>(g->reps.size())));
        if (rc) return rc;
        g->step_graphs.clear(); // captured with the previous fork structure
    }
    return FTS_OK;
}

int fts_batch_set_graph(fts_batch* b, int level)
{
    if (!b || level < 0 || level > 2) return fail(FTS_INVALID, "batch: graph level must be 0, 1 or 2");
    for (auto& g : b->groups) {
        if (g->stream) HIP_TRY(hipStreamSynchronize(g->stream));
        g->graph_level = level;
        g->step_graphs.clear();
    }
    return FTS_OK;
}

int fts_batch_set_timing(fts_batch* b, int enabled)
{
    if (!b) return fail(FTS_INVALID, "null batch");
    for (auto& g : b->groups) {
        g->timing = enabled != 0;
        g->tstat = fts_timing{};
    }
    return FTS_OK;
}

int fts_batch_get_timing(fts_batch* b, int group, fts_timing* out)
{
    if (!b || !out || group < 0 || group >= static_cast<int>(b->groups.size())) return fail(FTS_INVALID, "batch: bad argument");
    *out = b->groups[static_cast<size_t>(group)]->tstat;
    return FTS_OK;
}

} // extern "C"
