// Job-table farm: the in-process counterpart of the reference's MPI master/worker farm
// (src/mpidp.cpp).  The table format and option names are the reference's:
//   TITLE=<title>            (src/mpidp.cpp:212-217)
//   PARAM=<command line>     with $1, $2, ... replaced by the fields of a job line
//                            (src/mpidp.cpp:218-224, :787-834, replace_pattern :877-897)
//   <field 1>\t<field 2>...  one job per line
// and -tb table, -lg log, -rt retries, -ot k (field k names the output file; a retry
// appends ".<try>" to it, src/mpidp.cpp:803-810), -wl / -pg (accepted, no effect here);
// every other option is passed to each job in front of the expanded PARAM line.
//
// Instead of one process per job over MPI ranks, one process keeps the GPU busy:
//   prep threads (-pj, default min(8, cores/2)) do each job's CPU work: command line,
//     PDB reading, grid sizing, the receptor's shape/electrostatics grids (cached per
//     receptor file and grid, since a table repeats receptors);
//   the calling thread docks the prepared jobs on GPU 0 through libmdk contexts kept per
//     FFT size (re-created only when a job needs more ligand atoms, another t or another
//     batch; alive within a device-memory budget, -gm, the least recently used freed
//     first), taking first a prepared job whose size has a live context and differs from
//     the jobs in flight.  Two jobs are in flight while the next is set up: its receptor,
//     ligand and batches are queued on its own context before the GPU thread waits for the
//     oldest job's merge and writes its .out file, so the GPU runs the jobs' batches back
//     to back (a job of an in-flight job's size waits for that job, as they share a
//     context; the other one keeps the GPU busy meanwhile).
// Host memory stays bounded on long tables: at most two prepared jobs per preparation
// thread wait for the GPU, and the receptor grid cache keeps the grids of the last 16
// receptors (and of any receptor a waiting job still uses).
// A failed job is put back in the queue until it has run 1 + retries times
// (src/mpidp.cpp:372-543); the log ends with the per-job table (src/mpidp.cpp:900-935).
#include <sys/time.h>

#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <tuple>
#include <vector>

#include "mdk_host.h"

namespace mdk {
namespace {

double now_s() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

struct FarmOptions {
    std::string table, log = "./master.log";
    int out_option = 0, retries = 0, prep_threads = 0;
    double context_gib = 160.0;  // -gm: device memory for the live contexts (GiB)
    bool dry = false;
    std::vector<std::string> app;  // options passed to every job
};

// src/mpidp.cpp:148-191: farm options; the rest belongs to the application
int parse_options(int argc, char **argv, FarmOptions &o, std::string &err) {
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto value = [&](std::string &v) {
            if (i + 1 >= argc) {
                err = "option " + a + " needs a value";
                return false;
            }
            v = argv[++i];
            return true;
        };
        std::string v;
        if (a == "-tb") {
            if (!value(o.table)) return 1;
        } else if (a == "-lg") {
            if (!value(o.log)) return 1;
        } else if (a == "-ot") {
            if (!value(v)) return 1;
            o.out_option = std::atoi(v.c_str());
        } else if (a == "-rt") {
            if (!value(v)) return 1;
            o.retries = std::max(0, std::atoi(v.c_str()));
        } else if (a == "-wl" || a == "-pg") {
            if (!value(v)) return 1;
        } else if (a == "-pj") {
            if (!value(v)) return 1;
            o.prep_threads = std::atoi(v.c_str());
        } else if (a == "-gm") {  // not in the reference: this farm keeps a context per FFT size
            if (!value(v)) return 1;
            o.context_gib = std::atof(v.c_str());
            if (!(o.context_gib > 0)) {
                err = "option -gm needs a positive number of GiB";
                return 1;
            }
        } else if (a == "-dry") {
            o.dry = true;
        } else {
            o.app.push_back(a);
        }
    }
    if (o.table.empty()) {
        err = "[ERROR] no job table: -tb <table file>";
        return 1;
    }
    return 0;
}

// src/mpidp.cpp:251-275: spaces among the first ip characters are deleted
std::string erase_space(std::string s, size_t ip) {
    for (;;) {
        const size_t n = s.find(' ');
        if (n == std::string::npos || n >= ip) return s;
        s.erase(n, 1);
    }
}

// src/mpidp.cpp:877-897
std::string replace_pattern(const std::string &pattern, const std::string &position, const std::string &option) {
    std::string out;
    size_t before = 0, pos = 0;
    while ((pos = pattern.find(position, pos)) != std::string::npos) {
        out.append(pattern, before, pos - before);
        out.append(option);
        pos += position.size();
        before = pos;
    }
    out.append(pattern, before, std::string::npos);
    return out;
}

struct Table {
    std::string title = "MasterProc", param = "MPIDP";
    std::vector<std::string> jobs;
};

// src/mpidp.cpp:198-246 (an empty line is not a job here)
int read_table(const std::string &path, Table &t, std::string &err) {
    std::ifstream in(path);
    if (!in) {
        err = "[ERROR] Table file [" + path + "] was not opened!!";
        return 1;
    }
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        line = erase_space(line, 7);
        if (!strncmp(line.c_str(), "TITLE=", 6) || !strncmp(line.c_str(), "Title=", 6) ||
            !strncmp(line.c_str(), "title=", 6))
            t.title = line.substr(6);
        else if (!strncmp(line.c_str(), "PARAM=", 6) || !strncmp(line.c_str(), "Param=", 6) ||
                 !strncmp(line.c_str(), "param=", 6))
            t.param = line.substr(6);
        else if (!line.empty())
            t.jobs.push_back(line);
    }
    return 0;
}

std::vector<std::string> split(const std::string &s, char sep) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : s) {
        if (ch == sep) {
            out.push_back(cur);
            cur.clear();
        } else {
            cur += ch;
        }
    }
    out.push_back(cur);
    return out;
}

// The job's command line: the farm's other options, then PARAM with $i replaced by
// field i (src/mpidp.cpp:796-822; tokens split at spaces, empty tokens dropped).
std::vector<std::string> expand_job(const FarmOptions &o, const Table &t, const std::string &line, int attempt) {
    std::vector<std::string> fields = split(line, '\t');
    std::string p = t.param;
    for (size_t i = 0; i < fields.size(); i++) {
        std::string f = fields[i];
        if (o.out_option == (int)i + 1 && attempt > 0) f += "." + std::to_string(attempt);
        p = replace_pattern(p, "$" + std::to_string(i + 1), f);
    }
    std::vector<std::string> argv = {"megadock-hip"};
    argv.insert(argv.end(), o.app.begin(), o.app.end());
    for (const auto &tok : split(p, ' '))
        if (!tok.empty()) argv.push_back(tok);
    return argv;
}

// ------------------------------------------------------------------ prepared jobs
struct Grids {
    std::vector<float> r, i;
};

// Receptor grids shared by the jobs that dock the same receptor on the same grid.  An
// entry lives while a job holds it or while it is among the kKeep most recently used, so
// a long table of distinct receptors does not accumulate grids.
class GridCache {
public:
    using Key = std::tuple<std::string, int, float, float, float, float>;
    struct Entry {
        std::once_flag once;
        std::shared_ptr<const Grids> grids;
    };
    static constexpr size_t kKeep = 16;
    std::shared_ptr<Entry> get(const Key &k, const Receptor &rec, int ng, const Parameter &p) {
        std::shared_ptr<Entry> e;
        {
            std::lock_guard<std::mutex> g(mu_);
            e = map_[k].lock();
            if (!e) {
                e = std::make_shared<Entry>();
                map_[k] = e;
            }
            auto it = std::find(recent_.begin(), recent_.end(), e);
            if (it != recent_.end()) recent_.erase(it);
            recent_.push_front(e);
            if (recent_.size() > kKeep) recent_.pop_back();
            if (map_.size() > 4 * kKeep)  // forget the keys whose grids are gone
                for (auto m = map_.begin(); m != map_.end();) m = m->second.expired() ? map_.erase(m) : std::next(m);
        }
        std::call_once(e->once, [&] {
            auto gr = std::make_shared<Grids>();
            const size_t ng3 = (size_t)ng * ng * ng;
            gr->r.resize(ng3);
            gr->i.resize(ng3);
            std::vector<float> xyz(3 * (size_t)rec.num_atoms());
            rec.shift_center(xyz.data());
            rec.voxelize(ng, p.grid_width, grid_coord(ng, p.grid_width), xyz.data(), p.ace_ratio, p.rec_core,
                         p.elec_ratio, gr->r.data(), gr->i.data(), 1);  // the preparation threads run in parallel
            e->grids = gr;
        });
        return e;
    }
    size_t size() {
        std::lock_guard<std::mutex> g(mu_);
        size_t n = 0;
        for (auto &m : map_) n += !m.second.expired();
        return n;
    }

private:
    std::mutex mu_;
    std::map<Key, std::weak_ptr<Entry>> map_;
    std::deque<std::shared_ptr<Entry>> recent_;
};

struct Job {
    int id = 0;          // 1-based event number (src/mpidp.cpp:313)
    int attempt = 0;     // 0 for the first run, then 1..retries
    std::string line;
    // filled by preparation
    std::vector<std::string> argv;
    Parameter prm;
    std::unique_ptr<Receptor> rec;
    std::unique_ptr<Ligand> lig;
    int ng = 0;
    std::vector<float> gc;
    std::shared_ptr<GridCache::Entry> grids;  // grids->grids: the receptor grids
    mdk_ctx *ctx = nullptr;  // the context the job was issued on
    std::string err;
    double t_prep = 0, t_dock = 0;
    // filled by docking: the merged top-N, written to the .out file afterwards
    std::vector<SortScore> top;
};

struct JobRecord {
    std::string line, status = "not run", out_file, message;
    int tries = 0, ng = 0, rotations = 0;
    double seconds = 0;
};

std::mutex g_getopt_mu;  // Parameter::initialize uses getopt's global state

int prepare(Job &j, GridCache &cache) {
    const double t0 = now_s();
    {
        std::vector<char *> av;
        for (auto &a : j.argv) av.push_back(const_cast<char *>(a.c_str()));
        av.push_back(nullptr);
        std::lock_guard<std::mutex> g(g_getopt_mu);
        if (j.prm.initialize((int)j.argv.size(), av.data(), j.err)) return 1;
    }
    Parameter &p = j.prm;
    if (p.gpu_limit == 0) {
        j.err = "-G 0 asks for the CPU loop, which this build does not provide";
        return 1;
    }
    j.rec.reset(new Receptor(p.rec_file));
    j.lig.reset(new Ligand(p.lig_file));
    if (j.rec->initialize(p, j.err) || j.lig->initialize(p, j.err)) return 1;
    if (!p.num_fft_flag) {
        const int a = autogrid_receptor(*j.rec, p), b = autogrid_ligand(*j.lig, p);
        if (a < 0 || b < 0) {
            j.err = "grid sizing failed (protein larger than the FFT size table)";
            return 1;
        }
        j.ng = std::max(a, b);
    } else {
        if (check_grid(*j.rec, *j.lig, p, j.err)) return 1;
        j.ng = p.num_grid;
    }
    j.gc = grid_coord(j.ng, p.grid_width);
    j.grids = cache.get(GridCache::Key{p.rec_file, j.ng, p.grid_width, p.ace_ratio, p.rec_core, p.elec_ratio},
                        *j.rec, j.ng, p);
    j.t_prep = now_s() - t0;
    return 0;
}

// ------------------------------------------------------------------ device side
struct Slot {
    mdk_ctx *ctx = nullptr;
    int na = 0, t = 0, batch = 0;
    int rot_set = -1;  // rotation set loaded into the context
    long last_use = 0;
    double bytes = 0;  // device bytes (context_bytes)
};

// One context per FFT size, kept alive while their device buffers fit a budget (-gm, 160
// GiB of the GPU's 288 GB by default; context_bytes estimates each): a table
// with more sizes than fit frees the least recently used contexts (never one whose job is
// in flight) before it creates another.  Freeing a context waits for the
// whole device, so a context is made with room for ligands of up to kMinAtoms atoms and
// is not remade for a ligand that fits it.
class Contexts {
public:
    static constexpr int kMinAtoms = 16384;
    explicit Contexts(double budget_bytes) : budget_(budget_bytes) {}
    ~Contexts() {
        for (auto &kv : slots_)
            if (kv.second.ctx) mdk_destroy(kv.second.ctx);
    }
    // the context of the job's FFT size, able to hold its ligand, t and batch (its slot: the
    // context and the rotation set loaded into it); `busy` (the in-flight jobs' contexts,
    // another size) is never freed
    Slot *get(const Job &j, const std::vector<mdk_ctx *> &busy, std::string &err) {
        const int nf = 2 * j.ng, t = j.prm.num_sort;
        int batch = j.prm.batch > 0 ? j.prm.batch : auto_batch(nf, j.ng, t);
        batch = std::min(batch, std::max(1, j.prm.num_rot_angles));
        const int na = std::max(j.lig->num_atoms(), std::min(kMinAtoms, 65535));
        Slot *s = &slots_[nf];
        s->last_use = ++clock_;
        if (s->ctx && (s->na < j.lig->num_atoms() || s->t != t || s->batch < batch)) {
            mdk_destroy(s->ctx);
            destroyed++;
            *s = Slot();
        }
        if (!s->ctx) {
            const double need = context_bytes(nf, j.ng, t, batch, na);
            free_to(budget_ - need, nf, busy);
            int rc = mdk_create(0, nf, j.ng, na, t, batch, &s->ctx);
            if (rc && free_to(0, nf, busy) > 0) rc = mdk_create(0, nf, j.ng, na, t, batch, &s->ctx);  // out of memory?
            if (rc) {
                err = std::string("mdk_create: ") + mdk_last_error();
                *s = Slot();
                return nullptr;
            }
            created++;
            s->na = na;
            s->t = t;
            s->batch = batch;
            s->bytes = need;
            s->last_use = clock_;
        }
        return s;
    }
    bool has(int nf) const {
        auto it = slots_.find(nf);
        return it != slots_.end() && it->second.ctx != nullptr;
    }
    int created = 0, destroyed = 0;

private:
    // destroy the least recently used contexts other than size `keep_nf` and `busy` until
    // the others' bytes are at most `room`; returns how many were destroyed
    int free_to(double room, int keep_nf, const std::vector<mdk_ctx *> &busy) {
        int freed = 0;
        for (;;) {
            double live = 0;
            Slot *lru = nullptr;
            for (auto &kv : slots_)
                if (kv.first != keep_nf && kv.second.ctx) {
                    live += kv.second.bytes;
                    const bool in_use = std::find(busy.begin(), busy.end(), kv.second.ctx) != busy.end();
                    if (!in_use && (!lru || kv.second.last_use < lru->last_use)) lru = &kv.second;
                }
            if (live <= room || !lru) return freed;
            mdk_destroy(lru->ctx);
            destroyed++;
            *lru = Slot();
            freed++;
        }
    }
    std::map<int, Slot> slots_;
    long clock_ = 0;
    double budget_;
};

// The device part of a job in two calls: issue() uploads the job and queues its batches
// (mdk_dock_range returns while they run); complete() waits for the merged top-N, and for
// -O also scores and writes the detailed output, which needs the context.  The .out file
// is written by write_out(), from host data only, so the GPU thread can write one job's
// file while the next job's batches run.
int issue(Job &j, Contexts &cs, const std::vector<mdk_ctx *> &busy) {
    const double t0 = now_s();
    Parameter &p = j.prm;
    Slot *slot = cs.get(j, busy, j.err);
    if (!slot) return 1;
    mdk_ctx *ctx = j.ctx = slot->ctx;
    auto abi = [&](const char *what) {
        j.err = std::string(what) + ": " + mdk_last_error();
        return 1;
    };
    int &loaded = slot->rot_set;
    if (loaded != p.rotation_set) {  // the coefficients are made here, not kept by every waiting job
        std::vector<float> rotcoef(9 * (size_t)p.num_rot_angles);
        if (mdk_rotation_coef(p.zangle.data(), p.num_rot_angles, rotcoef.data())) return abi("mdk_rotation_coef");
        if (mdk_set_rotations(ctx, rotcoef.data(), p.num_rot_angles)) return abi("mdk_set_rotations");
        loaded = p.rotation_set;
    }
    const Grids &g = *j.grids->grids;
    if (mdk_set_receptor(ctx, g.r.data(), g.i.data())) return abi("mdk_set_receptor");
    const Ligand &L = *j.lig;
    if (mdk_set_ligand_atoms(ctx, L.num_atoms(), L.coords().data(), L.center(), L.radius.data(), L.charge.data(),
                             j.gc.data(), p.grid_width, p.lig_core))
        return abi("mdk_set_ligand_atoms");
    if (mdk_dock_range(ctx, 0, p.num_rot_angles)) return abi("mdk_dock_range");
    // the merge is queued right behind the job's last batch, before later jobs' work
    if (mdk_topn_start(ctx, p.num_rot_angles, p.num_output)) return abi("mdk_topn_start");
    j.t_dock = now_s() - t0;
    return 0;
}

// waits for the job's merged top-N, started by issue() on the context it used (still
// alive: Contexts never frees a busy context)
int complete(Job &j) {
    const double t0 = now_s();
    Parameter &p = j.prm;
    std::vector<mdk_sortscore> best(std::max(1, p.num_output));
    if (mdk_topn_finish(j.ctx, best.data())) {
        j.err = std::string("mdk_topn_finish: ") + mdk_last_error();
        return 1;
    }
    j.top.resize(best.size());
    std::memcpy(j.top.data(), best.data(), sizeof(SortScore) * best.size());
    if (p.detail_output_flag && write_detail(p, *j.rec, *j.lig, j.ng, j.top.data(), j.ctx, j.err)) return 1;
    j.t_dock += now_s() - t0;
    return 0;
}

int write_out(Job &j) {
    const double t0 = now_s();
    const int rc = write_decoys(j.prm, *j.rec, *j.lig, j.ng, j.top.data(), j.err);
    j.t_dock += now_s() - t0;
    return rc;
}

template <class T> class Queue {
public:
    explicit Queue(size_t cap = 0) : cap_(cap) {}
    // blocks while the queue holds cap items (cap 0: unbounded)
    void push(T v) {
        {
            std::unique_lock<std::mutex> g(mu_);
            space_.wait(g, [&] { return cap_ == 0 || q_.size() < cap_ || closed_; });
            q_.push_back(std::move(v));
            high_ = std::max(high_, q_.size());
        }
        cv_.notify_one();
    }
    // blocks until an item or close(); false when closed and empty
    bool pop(T &v) {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return !q_.empty() || closed_; });
        if (q_.empty()) return false;
        v = std::move(q_.front());
        q_.pop_front();
        space_.notify_one();
        return true;
    }
    // without waiting: false when the queue is empty, else as pop_best
    template <class R> bool try_pop_best(T &v, R rank) {
        std::lock_guard<std::mutex> g(mu_);
        if (q_.empty()) return false;
        take_best(v, rank);
        return true;
    }
    // as pop, but takes the item of lowest rank(item), the first of equal ones
    template <class R> bool pop_best(T &v, R rank) {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return !q_.empty() || closed_; });
        if (q_.empty()) return false;
        take_best(v, rank);
        return true;
    }
    void close() {
        {
            std::lock_guard<std::mutex> g(mu_);
            closed_ = true;
        }
        cv_.notify_all();
        space_.notify_all();
    }
    // the most items the queue has held at once
    size_t high_water() {
        std::lock_guard<std::mutex> g(mu_);
        return high_;
    }

private:
    template <class R> void take_best(T &v, R rank) {  // q_ not empty, mu_ held
        auto best = q_.begin();
        int rb = rank(*best);
        for (auto it = std::next(q_.begin()); it != q_.end() && rb > 0; ++it) {
            const int r = rank(*it);
            if (r < rb) {
                best = it;
                rb = r;
            }
        }
        v = std::move(*best);
        q_.erase(best);
        space_.notify_one();
    }
    size_t cap_, high_ = 0;
    std::mutex mu_;
    std::condition_variable cv_, space_;
    std::deque<T> q_;
    bool closed_ = false;
};

}  // namespace

int run_table(int argc, char **argv, std::string &err) {
    const double t_start = now_s();
    FarmOptions o;
    if (parse_options(argc, argv, o, err)) return 1;
    Table t;
    if (read_table(o.table, t, err)) return 1;
    std::ofstream log(o.log);
    if (!log) {
        err = "[ERROR] Log file [" + o.log + "] was not opened!!";
        return 1;
    }
    log << " MEGADOCK-compatible job farm on MI355X (libmdk)\n\n";
    log << "Table file    : -tb " << o.table << "\n";
    if (o.out_option) log << "Output option : -ot " << o.out_option << "\n";
    if (o.retries) log << "Retry option  : -rt " << o.retries << "\n";
    if (!o.app.empty()) {
        log << "Other options :";
        for (auto &a : o.app) log << " " << a;
        log << "\n";
    }
    log << "\nTITLE=" << t.title << "\nPARAM=" << t.param << "\n\n";
    const int njobs = (int)t.jobs.size();
    std::vector<JobRecord> rec(njobs);
    for (int i = 0; i < njobs; i++) rec[i].line = t.jobs[i];
    if (o.dry) {  // the command line of every job, nothing run
        for (int i = 0; i < njobs; i++) {
            char name[16];
            std::snprintf(name, sizeof name, "%05d", i + 1);
            log << name << "\t";
            const auto av = expand_job(o, t, t.jobs[i], 0);
            for (size_t k = 0; k < av.size(); k++) log << (k ? " " : "") << av[k];
            log << "\n";
        }
        return 0;
    }
    int ndev = 0;
    if (mdk_device_count(&ndev) || ndev < 1) {
        err = "no GPU visible: this build runs the rotation loop on the GPU";
        return 1;
    }
    int nprep = o.prep_threads;
    if (nprep < 1) nprep = std::max(1, std::min(8, (int)std::thread::hardware_concurrency() / 2));

    const int wait_limit = 2 * nprep;  // prepared jobs wait in a bounded queue
    Queue<std::unique_ptr<Job>> todo, ready((size_t)wait_limit);
    std::mutex rec_mu;
    int outstanding = njobs;  // jobs neither finished nor given up (guarded by rec_mu)
    for (int i = 0; i < njobs; i++) {
        std::unique_ptr<Job> j(new Job());
        j->id = i + 1;
        j->line = t.jobs[i];
        todo.push(std::move(j));
    }
    if (njobs == 0) todo.close();
    // a job that failed: back into the queue while it has tries left
    auto finish = [&](std::unique_ptr<Job> j, bool ok) {
        std::lock_guard<std::mutex> g(rec_mu);
        JobRecord &r = rec[j->id - 1];
        r.tries = j->attempt + 1;
        r.ng = j->ng;
        r.rotations = j->prm.num_rot_angles;
        r.seconds += j->t_prep + j->t_dock;
        r.out_file = j->prm.out_file;
        if (ok) {
            r.status = "done";
            r.message.clear();
        } else {
            r.status = "failed";
            r.message = j->err;
            if (j->attempt < o.retries) {
                std::unique_ptr<Job> again(new Job());
                again->id = j->id;
                again->line = j->line;
                again->attempt = j->attempt + 1;
                todo.push(std::move(again));
                return;
            }
        }
        if (--outstanding == 0) todo.close();
    };
    GridCache cache;
    std::vector<std::thread> workers;
    std::mutex ready_mu;
    int prep_alive = nprep;
    for (int w = 0; w < nprep; w++)
        workers.emplace_back([&] {
            std::unique_ptr<Job> j;
            while (todo.pop(j)) {
                j->argv = expand_job(o, t, j->line, j->attempt);
                int rc = 1;
                try {
                    rc = prepare(*j, cache);
                } catch (const std::exception &e) {  // e.g. out of host memory: this job fails
                    j->err = std::string("preparation: ") + e.what();
                }
                if (rc) {
                    j->t_prep = 0;
                    finish(std::move(j), false);
                } else {
                    ready.push(std::move(j));
                }
            }
            std::lock_guard<std::mutex> g(ready_mu);
            if (--prep_alive == 0) ready.close();
        });
    int drain_empty = 0, drain_same = 0, ctx_made = 0, ctx_freed = 0;  // the GPU's waits between jobs
    {
        Contexts cs(o.context_gib * 1024 * 1024 * 1024);
        // in flight: issued jobs whose merge is not yet waited for, oldest first, each on
        // the context of its own FFT size; written: merged, their .out files still to write
        constexpr size_t kDepth = 2;  // jobs queued on the GPU while the farm waits for a merge
        std::deque<std::unique_ptr<Job>> inflight;
        std::vector<std::unique_ptr<Job>> written;
        std::unique_ptr<Job> j;
        auto report = [&](std::unique_ptr<Job> x, bool ok) {
            std::printf("# job %05d (%s): %s\n", x->id, x->line.c_str(), ok ? "done" : x->err.c_str());
            std::fflush(stdout);
            finish(std::move(x), ok);
        };
        auto write_pending = [&] {
            for (auto &w : written) {
                bool ok = false;
                try {
                    ok = write_out(*w) == 0;
                } catch (const std::exception &e) {
                    w->err = std::string("output: ") + e.what();
                }
                report(std::move(w), ok);
            }
            written.clear();
        };
        auto complete_oldest = [&] {
            std::unique_ptr<Job> x = std::move(inflight.front());
            inflight.pop_front();
            bool ok = false;
            try {
                ok = complete(*x) == 0;
            } catch (const std::exception &e) {
                x->err = std::string("docking: ") + e.what();
            }
            write_pending();  // the jobs merged before, if any
            if (ok) written.push_back(std::move(x));
            else report(std::move(x), false);
        };
        auto in_flight = [&](int ng) {
            for (auto &x : inflight)
                if (x->ng == ng) return true;
            return false;
        };
        // the prepared job to take: first one of a size no job in flight has (its batches
        // then queue behind theirs), best with a live context (fewer contexts made and freed
        // when the table mixes more sizes than stay alive); one of an in-flight job's size
        // only when no other is prepared
        auto rank = [&](const std::unique_ptr<Job> &x) {
            return in_flight(x->ng) ? 2 : cs.has(2 * x->ng) ? 0 : 1;
        };
        for (;;) {
            if (!ready.try_pop_best(j, rank)) {
                if (!inflight.empty()) drain_empty++;
                while (!inflight.empty()) complete_oldest();  // nothing prepared: finish what is queued before waiting
                write_pending();
                if (!ready.pop_best(j, rank)) break;
            }
            // a job of an in-flight job's size shares its context: that job (and the ones
            // before it) complete first; the GPU drains only if it was the newest
            if (in_flight(j->ng)) {
                while (in_flight(j->ng)) complete_oldest();
                if (inflight.empty()) drain_same++;
            }
            std::vector<mdk_ctx *> busy;
            for (auto &x : inflight) busy.push_back(x->ctx);
            bool ok = false;
            try {
                ok = issue(*j, cs, busy) == 0;  // queued behind the jobs in flight
            } catch (const std::exception &e) {
                j->err = std::string("docking: ") + e.what();
            }
            // wait for the oldest merge while the newer jobs' batches run; its .out file is
            // written then too
            while (inflight.size() >= kDepth) complete_oldest();
            write_pending();
            if (ok) inflight.push_back(std::move(j));
            else report(std::move(j), false);
        }
        while (!inflight.empty()) complete_oldest();
        write_pending();
        ctx_made = cs.created;
        ctx_freed = cs.destroyed;
    }
    for (auto &w : workers) w.join();

    // per-job report (src/mpidp.cpp:900-935)
    int failed = 0;
    long long rotations = 0;
    log << "JOB\tTRIES\tSTATUS\tNG\tROTATIONS\tSECONDS\tOUTPUT\tLINE\n";
    for (int i = 0; i < njobs; i++) {
        const JobRecord &r = rec[i];
        char name[16];
        std::snprintf(name, sizeof name, "%05d", i + 1);
        log << name << "\t" << r.tries << "\t" << r.status << "\t" << r.ng << "\t" << r.rotations << "\t" << r.seconds
            << "\t" << r.out_file << "\t" << r.line;
        if (!r.message.empty()) log << "\t# " << r.message;
        log << "\n";
        if (r.status != "done") failed++;
        else rotations += r.rotations;
    }
    const double el = now_s() - t_start;
    log << "\nPrepared jobs waiting for the GPU: at most " << ready.high_water() << " (limit " << wait_limit << ")";
    log << "\nGPU drained between jobs: " << drain_empty << " with no job prepared, " << drain_same
        << " for a job of the running job's FFT size; contexts made " << ctx_made << ", freed " << ctx_freed;
    log << "\nJobs done = " << njobs - failed << " / " << njobs << "\nElapsed time  = " << el << " sec.\n"
        << "Rotations docked = " << rotations << " (" << (el > 0 ? rotations / el : 0.0) << " per second)\n";
    if (failed) {
        err = std::to_string(failed) + " of " + std::to_string(njobs) + " jobs failed (see " + o.log + ")";
        return 1;
    }
    return 0;
}

}  // namespace mdk
