// hermite_run: the reference's command-line driver on the gfx950 back end.
//
// Reads the reference's parameter tag file on stdin (Parameter::read,
// SRC/parameter.cpp:5-55: one "tag value" per line, lines not starting with
// a letter ignored, dtmaxinv / epsinv give reciprocals), prints it
// (Parameter::print, :57-80) and runs the main loop of SRC/compile.cpp:127-216:
// start from a masaki snapshot (init_step) or a dump (restore +
// reinitialize_steps), integrate dt_max intervals to tend, and at the log,
// snapshot and dump intervals print the print_stat line, write
// snapout_base / dumpout_base files (printf patterns of the id) and print the
// profile (Profile::show, SRC/nbodysystem.h:122-159).  Output lines and file
// formats are the reference's.
//
// Tags beyond the reference's (the reference fixes these at compile time or
// has no counterpart):
//   variant fp64|split|exact   force kernel (default fp64)
//   device  <ordinal>          HIP device (default 0)
//   dtmin <dt> | dtmininv <1/dt>   lower clamp on the timestep (default none)
//   resident 0|1               1: integrate with the GPU-resident block step
//                              (hhr_*) between output times instead of the
//                              host driver (orders 4, 6, 8; default 0).  The
//                              outputs keep their formats; the particles are
//                              not permuted between block steps, so the j
//                              summation order, and with it the last bits of
//                              a run, differ from the host driver's.
// The order comes from the "order" tag (4, 6 or 8; "8dd" selects the
// double-double 8th-order build, which the reference selects at compile
// time with -DEIGHTHDD and whose parser reads "8dd" as 8); the reference
// binary asserts the order against the one it was compiled for.
//
//   hermite_run < inp.6th          run
//   hermite_run --params < inp     parse and print the parameters only
#include "hermite_driver.h"
#include "hermite_hip.h"

#include <omp.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Params {
    std::string snapin, snapout_base, dumpin, dumpout_base;
    int order = 0, snapid = 0, dumpid = 0;
    bool dd = false;  // "order 8dd" 
    double tend, eta, eta_s, dtmax, eps, kt_for_eps;
    double log_interval, snap_interval, dump_interval;
    std::string variant = "fp64";
    int device = 0;
    int resident = 0;
    double dtmin = 0.0;

    Params()
    {
        // unset values are the NaN of 0.0/0.0 (sign bit set on x86, printed
        // "-nan"), as in SRC/parameter.h:33
        const double nan = std::copysign(std::numeric_limits<double>::quiet_NaN(), -1.0);
        const double inf = std::numeric_limits<double>::infinity();
        tend = eta = eta_s = dtmax = eps = kt_for_eps = nan;
        log_interval = snap_interval = dump_interval = inf;
    }

    // Tag parser with the reference's rules (SRC/parameter.cpp:5-55).
    void read(std::istream &is)
    {
        for (std::string line; std::getline(is, line);) {
            std::istringstream ss(line);
            std::string tag;
            ss >> tag;
            if (ss.fail() || !std::isalpha((unsigned char)tag[0])) continue;
            auto inv = [&](double &x) { double t; ss >> t; x = 1.0 / t; };
            if (tag == "snapin") ss >> snapin;
            else if (tag == "snapout_base") ss >> snapout_base;
            else if (tag == "dumpin") ss >> dumpin;
            else if (tag == "dumpout_base") ss >> dumpout_base;
            else if (tag == "order") {
                std::string v;
                ss >> v;
                order = std::atoi(v.c_str());
                dd = v == "8dd";
            }
            else if (tag == "snapid") ss >> snapid;
            else if (tag == "dumpid") ss >> dumpid;
            else if (tag == "tend") ss >> tend;
            else if (tag == "eta") ss >> eta;
            else if (tag == "eta_s") ss >> eta_s;
            else if (tag == "dtmax") ss >> dtmax;
            else if (tag == "dtmaxinv") inv(dtmax);
            else if (tag == "eps") ss >> eps;
            else if (tag == "epsinv") inv(eps);
            else if (tag == "kt_for_eps") ss >> kt_for_eps;
            else if (tag == "log_interval") ss >> log_interval;
            else if (tag == "snap_interval") ss >> snap_interval;
            else if (tag == "dump_interval") ss >> dump_interval;
            else if (tag == "variant") ss >> variant;
            else if (tag == "device") ss >> device;
            else if (tag == "dtmin") ss >> dtmin;
            else if (tag == "resident") ss >> resident;
            else if (tag == "dtmininv") inv(dtmin);
            else std::cerr << "unmatched tag : " << tag << std::endl;
        }
    }

    // Parameter::print (SRC/parameter.cpp:57-80), then the extra tags.
    void print(std::ostream &os) const
    {
        os << "snapin : " << snapin << "\n" << "snapout_base : " << snapout_base << "\n"
           << "dumpin : " << dumpin << "\n" << "dumpout_base : " << dumpout_base << "\n"
           << "order : " << order << "\n" << "snapid : " << snapid << "\n" << "dumpid : " << dumpid << "\n"
           << "tend : " << tend << "\n" << "eta : " << eta << "\n" << "eta_s : " << eta_s << "\n"
           << "dtmax : " << dtmax << "\n" << "eps : " << eps << "\n" << "kt_for_eps : " << kt_for_eps << "\n"
           << "log_interval : " << log_interval << "\n" << "snap_interval : " << snap_interval << "\n"
           << "dump_interval : " << dump_interval << std::endl;
    }

    void print_extra(std::ostream &os) const
    {
        os << "variant : " << variant << "\n" << "device : " << device << "\n" << "dtmin : " << dtmin << "\n"
           << "resident : " << resident << std::endl;
    }

    // Parameter::assert_check (SRC/parameter.h:56-65); returns the failed
    // condition or an empty string.
    std::string check() const
    {
        if (!(snapin.length() > 0 || dumpin.length() > 0)) return "snapin or dumpin must be given";
        if (!(order == 4 || order == 6 || order == 8)) return "order must be 4, 6 or 8";
        if (!(tend > 0)) return "tend > 0";
        if (!(eta > 0)) return "eta > 0";
        if (!(eta_s > 0)) return "eta_s > 0";
        if (!(dtmax > 0)) return "dtmax > 0";
        if (!(std::log2(dtmax) == int(std::log2(dtmax)))) return "dtmax must be a power of two";
        if (!(eps >= 0.0 || kt_for_eps >= 0.0)) return "eps >= 0 or kt_for_eps >= 0";
        if (variant != "fp64" && variant != "split" && variant != "exact") return "variant must be fp64, split or exact";
        if (!(dtmin >= 0.0)) return "dtmin >= 0";
        if (dd && variant == "split") return "order 8dd has no split variant";
        if (resident != 0 && resident != 1) return "resident must be 0 or 1";
        if (resident && dd) return "order 8dd has no GPU-resident driver";
        return "";
    }
};

// printf-style file name from a pattern with one integer conversion
// (%d, %i or %u with optional flags and width), as the reference's
// snprintf(fname, 256, base, id) uses it; anything else is rejected.
bool format_name(const std::string &pat, int id, std::string &out)
{
    out.clear();
    bool used = false;
    for (size_t i = 0; i < pat.size(); i++) {
        if (pat[i] != '%') {
            out += pat[i];
            continue;
        }
        if (i + 1 < pat.size() && pat[i + 1] == '%') {
            out += '%';
            i++;
            continue;
        }
        size_t j = i + 1;
        while (j < pat.size() && std::strchr("-+ 0#", pat[j])) j++;
        while (j < pat.size() && std::isdigit((unsigned char)pat[j])) j++;
        if (used || j >= pat.size() || !std::strchr("diu", pat[j])) return false;
        std::string spec = pat.substr(i, j - i) + "d";
        char buf[64];
        std::snprintf(buf, sizeof buf, spec.c_str(), id);
        out += buf;
        used = true;
        i = j;
    }
    return out.size() < 256;
}

int variant_code(const std::string &v) { return v == "split" ? HH_SPLIT : v == "exact" ? HH_EXACT : HH_FP64; }
int pwords(int o) { return o == 4 ? 24 : o == 6 ? 28 : o == 8 ? 36 : 40; }
int flops(int o) { return o == 4 ? 60 : o == 6 ? 97 : 144; }

// Profile of the reference (SRC/nbodysystem.h:57-194) over the driver's
// phase clocks: times since the last flush, TOTAL from the main loop.
struct Prof {
    enum { NPH = 9, TOTAL = 8 };
    double base[NPH] = {0}, total = 0.0, tbeg = 0.0;
    long ns_prev = 0, nbs_prev = 0;
    bool resident = false;  // no host predict phase: no bandwidth line
    void flush(hhd_sys *s)
    {
        hhd_profile(s, base);
        total = 0.0;
    }
    void beg() { tbeg = omp_get_wtime(); }
    void end() { total += omp_get_wtime() - tbeg; }
    void show(hhd_sys *s, long nbody, int order)
    {
        static const char *names[NPH + 1] = {"force   ", "pot     ", "predict ", "correct ", "sort    ",
                                             "scan    ", "set_jp  ", "I/O     ", "total   ", "misc    "};
        double now[NPH], t[NPH + 1];
        hhd_profile(s, now);
        for (int i = 0; i < TOTAL; i++) t[i] = now[i] - base[i];
        t[TOTAL] = total;
        t[NPH] = total;
        for (int i = 0; i < TOTAL; i++) t[NPH] -= t[i];
        const long ns = hhd_num_step_tot(s) - ns_prev, nbs = hhd_num_bstep_tot(s) - nbs_prev;
        ns_prev = hhd_num_step_tot(s);
        nbs_prev = hhd_num_bstep_tot(s);
        for (int i = 0; i <= NPH; i++)
            std::fprintf(stderr, " %s : %e sec, %e usec : %6.2f %%\n", names[i], t[i], t[i] * (1.e6 / nbs),
                         100.0 * t[i] / t[TOTAL]);
        const double nact = double(ns) / double(nbs);
        const double gflops = (((1.0e-9 * ns) * nbody) * flops(order)) / total;
        std::fprintf(stdout, "## nbody  wtime(sec)  Gflops  wtime/block(usec)  nact\n");
        std::fprintf(stdout, "## %ld %e %e %e %e\n", nbody, total, gflops, total / nbs * 1.e6, nact);
        const double nread = double(nbs) * nbody * 8.0 * hh_jp_words(order);
        const double nwrite = double(nbs) * nbody * 8.0 * hh_pred_words(order);
        if (!resident) std::fprintf(stdout, "prediction bandwidth = %f GB/s\n", (nread + nwrite) / t[2] * 1.e-9);
        std::fflush(stderr);
        std::fflush(stdout);
    }
};

// print_stat (SRC/nbodysystem.h:934-969): stderr block and the stdout row.
void print_stat(hhd_sys *s, long nbody, int order, double wtime)
{
    double st[9];
    if (hhd_stat(s, st) != HH_OK) return;
    const double t0 = omp_get_wtime();
    const long nbs = (long)st[1], ns = (long)st[2], nbs_tot = (long)st[3], ns_tot = (long)st[4];
    const double gflops = (((1.0e-9 * ns) * nbody) * flops(order)) / wtime;
    std::fprintf(stderr, "t = %f\n", st[0]);
    std::fprintf(stderr, " steps: %ld %ld %ld %ld\n", nbs, ns, nbs_tot, ns_tot);
    std::fprintf(stderr, " nact : %f\n", st[5]);
    std::fprintf(stderr, " de(local/global) : %+e %+e\n", st[6], st[7]);
    std::fprintf(stderr, " %f sec, %f Gflops\n", wtime, gflops);
    std::fprintf(stdout, "%f %6ld %6ld %10ld %10ld  %8.2f  %+e  %+e  %f  %f \n", st[0], nbs, ns, nbs_tot, ns_tot, st[5],
                 st[6], st[7], wtime, gflops);
    std::fflush(stdout);
    hhd_add_io_time(s, omp_get_wtime() - t0);
}

int die(hhd_sys *s, const std::string &what)
{
    std::fprintf(stderr, "hermite_run: %s%s%s\n", what.c_str(), s ? ": " : "", s ? hhd_error(s) : "");
    if (s) hhd_destroy(s);
    return 1;
}

}  // namespace

int main(int argc, char **argv)
{
    Params prm;
    prm.read(std::cin);
    prm.print(std::cout);
    prm.print_extra(std::cerr);  // stdout keeps the reference's lines
    if (argc > 1 && std::strcmp(argv[1], "--params") == 0) return 0;
    const std::string bad = prm.check();
    if (!bad.empty()) return die(nullptr, "parameter check failed: " + bad);

    const int order = prm.dd ? HH_ORDER_8DD : prm.order;
    int snapid = prm.snapid, dumpid = prm.dumpid;
    const double eps2 = prm.eps * prm.eps;
    hhd_sys *s = nullptr;
    long n = 0;
    double tsys = 0.0;
    if (prm.snapin.length() > 0) {
        if (hh_read_snapshot(prm.snapin.c_str(), order, &snapid, &n, &tsys, nullptr) != HH_OK)
            return die(nullptr, hh_io_error());
        std::vector<double> recs((size_t)n * pwords(order));
        if (hh_read_snapshot(prm.snapin.c_str(), order, &snapid, &n, &tsys, recs.data()) != HH_OK)
            return die(nullptr, hh_io_error());
        std::fprintf(stderr, "read snapshot (form M.I.), n = %ld, t = %f\n", n, tsys);
        std::fprintf(stderr, "allocate resources for n = %ld\n", n);
        std::fprintf(stderr, "NUM_THREADS = %d\n", omp_get_max_threads());
        s = hhd_create(order, (int)n, recs.data(), tsys, eps2, prm.eta, prm.eta_s, prm.dtmax, prm.dtmin,
                       variant_code(prm.variant), prm.device);
        if (!s || hhd_status(s) != HH_OK) return die(s, "cannot create the driver");
        if (prm.kt_for_eps >= 0.0) {
            const double eps = 4.0 / n / prm.kt_for_eps;
            std::fprintf(stderr, "eps is set to %e\n", eps);
            hhd_set_eps2(s, eps * eps);
        }
        hhd_set_verbose(s, 1);
        if (hhd_init_step(s) != HH_OK) return die(s, "init_step");
    } else {
        double e0, eprev;
        long nstep_tot, nbstep_tot;
        if (hh_read_dump(prm.dumpin.c_str(), order, &dumpid, &n, &tsys, &e0, &eprev, &nstep_tot, &nbstep_tot,
                         nullptr) != HH_OK)
            return die(nullptr, hh_io_error());
        std::vector<double> recs((size_t)n * pwords(order));
        if (hh_read_dump(prm.dumpin.c_str(), order, &dumpid, &n, &tsys, &e0, &eprev, &nstep_tot, &nbstep_tot,
                         recs.data()) != HH_OK)
            return die(nullptr, hh_io_error());
        std::fprintf(stderr, "read dump file : <%s>, order=%d, nbody=%ld, tsys=%f\n", prm.dumpin.c_str(), order, n,
                     tsys);
        std::fprintf(stderr, "allocate resources for n = %ld\n", n);
        std::fprintf(stderr, "NUM_THREADS = %d\n", omp_get_max_threads());
        s = hhd_create(order, (int)n, recs.data(), tsys, eps2, prm.eta, prm.eta_s, prm.dtmax, prm.dtmin,
                       variant_code(prm.variant), prm.device);
        if (!s || hhd_status(s) != HH_OK) return die(s, "cannot create the driver");
        hhd_set_restart_state(s, e0, eprev, nstep_tot, nbstep_tot);
        if (prm.kt_for_eps >= 0.0) {
            const double eps = 4.0 / n / prm.kt_for_eps;
            std::fprintf(stderr, "eps is set to %e\n", eps);
            hhd_set_eps2(s, eps * eps);
        }
        if (hhd_reinitialize_steps(s, prm.eta, prm.dtmax) != HH_OK) return die(s, "reinitialize_steps");
    }

    // GPU-resident mode: the block steps run in hhr_* from the state the
    // host driver has initialised; the records come back at each output time
    hhr_ctx *r = nullptr;
    std::vector<double> rbuf;
    long ns0 = 0, nbs0 = 0, hits0 = 0;
    if (prm.resident) {
        rbuf.resize((size_t)n * pwords(order));
        hhd_get_ptcl(s, rbuf.data());
        const double par[5] = {hhd_eps2(s), prm.eta, prm.dtmax, prm.dtmin, hhd_tsys(s)};
        const int nn = (int)n;
        if (hhr_create(order, variant_code(prm.variant), prm.device, 1, &nn, rbuf.data(), par, &r) != HH_OK)
            return die(s, std::string("hhr_create: ") + hhr_last_error(nullptr));
        ns0 = hhd_num_step(s);
        nbs0 = hhd_num_bstep(s);
        hits0 = hhd_dtmin_hits(s);
        std::fprintf(stderr, "GPU-resident block steps\n");
    }
    auto fail_run = [&](const std::string &what) {
        std::string msg = what;
        if (r) {
            const char *e = hhr_last_error(r);
            if (e && *e) msg += std::string(": ") + e;
            hhr_destroy(r);
        }
        return die(s, msg);
    };

    double tnext_log = hhd_tsys(s) + prm.log_interval;
    double tnext_snap = hhd_tsys(s) + prm.snap_interval;
    double tnext_dump = hhd_tsys(s) + prm.dump_interval;
    double wt_log = omp_get_wtime();
    Prof prof;
    prof.resident = r != nullptr;
    prof.ns_prev = hhd_num_step_tot(s);
    prof.nbs_prev = hhd_num_bstep_tot(s);
    prof.flush(s);
    prof.beg();
    while (hhd_tsys(s) < prm.tend) {
        if (r) {
            // whole dt_max intervals up to the next output time
            const double t0 = hhd_tsys(s);
            const double tev = std::min(std::min(tnext_log, tnext_snap), std::min(tnext_dump, prm.tend));
            const double target = t0 + std::max(1.0, std::ceil((tev - t0) / prm.dtmax)) * prm.dtmax;
            if (hhr_run(r, target, 1, 64) != HH_OK) return fail_run("hhr_run");
            long long ns = 0, nbs = 0, hits = 0;
            double tr = 0.0;
            if (hhr_stats(r, 0, &ns, &nbs, &tr, &hits) != HH_OK || hhr_get_particles(r, 0, rbuf.data()) != HH_OK)
                return fail_run("hhr_stats");
            if (hhd_load_state(s, rbuf.data(), tr, ns0 + (long)ns, nbs0 + (long)nbs, hits0 + (long)hits) != HH_OK)
                return fail_run("hhd_load_state");
        } else if (hhd_integrate_one_dtmax(s) != HH_OK) {
            return die(s, "integrate_one_dtmax");
        }
        const double t = hhd_tsys(s);
        if (t >= tnext_log) {
            tnext_log += prm.log_interval;
            const double wt = omp_get_wtime();
            print_stat(s, n, order, wt - wt_log);
            wt_log = wt;
        }
        if (t >= tnext_snap) {
            tnext_snap += prm.snap_interval;
            snapid++;
            if (prm.snapout_base.length() > 0) {
                std::string fname;
                if (!format_name(prm.snapout_base, snapid, fname)) return fail_run("bad snapout_base pattern");
                std::fprintf(stderr, "snap file : %s\n", fname.c_str());
                if (hhd_write_snapshot(s, fname.c_str(), snapid) != HH_OK) return fail_run(hh_io_error());
            }
        }
        if (t >= tnext_dump) {
            tnext_dump += prm.dump_interval;
            dumpid++;
            prof.end();
            prof.show(s, n, order);
            if (prm.dumpout_base.length() > 0) {
                std::string fname;
                if (!format_name(prm.dumpout_base, dumpid, fname)) return fail_run("bad dumpout_base pattern");
                std::fprintf(stderr, "dump file : %s\n", fname.c_str());
                if (hhd_dump(s, fname.c_str(), dumpid) != HH_OK) return fail_run(hh_io_error());
            }
            prof.flush(s);
            prof.beg();
        }
    }
    if (r) hhr_destroy(r);
    hhd_destroy(s);
    return 0;
}
