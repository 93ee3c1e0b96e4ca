/* Threaded single-call driver: host threads call the reference-signature C++ layer
 * (gwh::MCMC_likelihood_extrinsic, include/gwat_hip_compat.hpp) concurrently, the way the reference's
 * samplers call MCMC_likelihood_extrinsic from their worker threads (mcmc_sampler.cpp:4585-4600).  The engine
 * coalesces the concurrent calls into shared launches; this program measures the calls per second and writes
 * every result so that a caller can compare them bit for bit with one-thread runs.
 *
 *   gwh_threaded_calls <input> <output> <threads> <content|pointer|register> <method> [repeats]
 *
 * input (native byte order): int32 n, int32 ndet, int32 det[ndet], int32 count, then f[n], per detector
 * data[n] (complex, re/im interleaved) and psd[n], then `count` gwh_params records.  Thread t evaluates the
 * records t, t + threads, ... `repeats` times (results of the last pass are written: count doubles).
 * One line of JSON on stdout: threads, calls, seconds, calls_per_s, and the coalescing counters. */
#include <atomic>
#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gwat_hip_compat.hpp"

namespace {

/* the field names of gen_params_base<double> that the C++ layer reads */
struct SamplerParams {
    double mass1 = 0, mass2 = 0, Luminosity_Distance = 0, spin1[3] = {0, 0, 0}, spin2[3] = {0, 0, 0}, tc = 0,
           phiRef = 0, f_ref = 0, incl_angle = 0, RA = 0, DEC = 0, psi = 0, gmst = 0, chip = -1, phip = -1;
    bool shift_time = true, shift_phase = true, sky_average = false, dep_postmerger = false,
         equatorial_orientation = false, horizon_coord = false, NSflag1 = false, NSflag2 = false;
    int Nmod = 0, Nmod_phi = 0, Nmod_sigma = 0, Nmod_beta = 0, Nmod_alpha = 0;
    double *bppe = nullptr, *betappe = nullptr;
    int *phii = nullptr, *sigmai = nullptr, *betai = nullptr, *alphai = nullptr;
    double *delta_phi = nullptr, *delta_sigma = nullptr, *delta_beta = nullptr, *delta_alpha = nullptr;
};

SamplerParams from_record(const gwh_params &p)
{
    SamplerParams g;
    g.mass1 = p.mass1;
    g.mass2 = p.mass2;
    g.Luminosity_Distance = p.DL;
    for (int i = 0; i < 3; i++) {
        g.spin1[i] = p.spin1[i];
        g.spin2[i] = p.spin2[i];
    }
    g.tc = p.tc;
    g.phiRef = p.phiRef;
    g.f_ref = p.f_ref;
    g.incl_angle = p.incl;
    g.RA = p.RA;
    g.DEC = p.DEC;
    g.psi = p.psi;
    g.gmst = p.gmst;
    g.chip = p.chip;
    g.phip = p.phip;
    g.shift_time = p.shift_time != 0;
    g.shift_phase = p.shift_phase != 0;
    g.dep_postmerger = p.dep_postmerger != 0;
    return g;
}

bool read_all(FILE *fp, void *dst, size_t bytes) { return fread(dst, 1, bytes, fp) == bytes; }

int fail(const char *msg)
{
    fprintf(stderr, "gwh_threaded_calls: %s\n", msg);
    return 2;
}

} // namespace

int main(int argc, char **argv)
{
    if (argc < 6) return fail("usage: <input> <output> <threads> <content|pointer|register> <method> [repeats]");
    const int threads = atoi(argv[3]);
    const std::string mode = argv[4], method = argv[5];
    const int repeats = argc > 6 ? atoi(argv[6]) : 1;
    if (threads < 1 || threads > 1024 || repeats < 1) return fail("bad thread or repeat count");
    FILE *fp = fopen(argv[1], "rb");
    if (!fp) return fail("cannot open the input");
    int n = 0, ndet = 0, count = 0;
    int det_ids[4];
    bool ok = read_all(fp, &n, 4) && read_all(fp, &ndet, 4) && ndet >= 1 && ndet <= 4 && n >= 2 &&
              read_all(fp, det_ids, 4 * (size_t)ndet) && read_all(fp, &count, 4) && count >= 1;
    if (!ok) return fail("bad input header");
    std::vector<double> f(n);
    std::vector<std::vector<std::complex<double>>> data(ndet, std::vector<std::complex<double>>(n));
    std::vector<std::vector<double>> psd(ndet, std::vector<double>(n));
    ok = read_all(fp, f.data(), 8 * (size_t)n);
    for (int d = 0; d < ndet && ok; d++)
        ok = read_all(fp, data[d].data(), 16 * (size_t)n) && read_all(fp, psd[d].data(), 8 * (size_t)n);
    std::vector<gwh_params> rec(count);
    ok = ok && read_all(fp, rec.data(), sizeof(gwh_params) * (size_t)count);
    fclose(fp);
    if (!ok) return fail("short input");

    /* the sampler's arrays, as the reference keeps them: one frequency array per detector (shared here) */
    std::vector<double *> fr(ndet, f.data()), ps(ndet);
    std::vector<std::complex<double> *> da(ndet);
    std::vector<int> lens(ndet, n);
    std::vector<std::string> names(ndet);
    static const char *det_names[] = {"Hanford", "Livingston", "Virgo", "Kagra", "Indigo", "CE", "ET1", "ET2",
                                      "ET3"};
    for (int d = 0; d < ndet; d++) {
        ps[d] = psd[d].data();
        da[d] = data[d].data();
        if (det_ids[d] < 0 || det_ids[d] >= (int)(sizeof(det_names) / sizeof(det_names[0])))
            return fail("detector id out of range");
        names[d] = det_names[det_ids[d]];
        if (gwh_detector_id(names[d].c_str()) != det_ids[d]) return fail("detector name table mismatch");
    }
    if (gwh_init(0) != GWH_OK) return fail(gwh_last_error());
    int handle = -1;
    if (mode == "pointer") {
        if (gwh_set_single_call_cache(GWH_CACHE_BY_POINTER) != GWH_OK) return fail(gwh_last_error());
    } else if (mode == "register") {
        std::vector<const double *> cf(fr.begin(), fr.end()), cp(ps.begin(), ps.end()), cd(ndet);
        for (int d = 0; d < ndet; d++) cd[d] = reinterpret_cast<const double *>(da[d]);
        if (gwh_host_data_register(ndet, lens.data(), cf.data(), cd.data(), cp.data(), nullptr, &handle) != GWH_OK)
            return fail(gwh_last_error());
    } else if (mode != "content") {
        return fail("mode is content, pointer or register");
    }

    std::vector<SamplerParams> gp(count);
    for (int i = 0; i < count; i++) gp[i] = from_record(rec[i]);
    std::vector<double> out(count, 0.0);
    auto call = [&](int i) {
        return gwh::MCMC_likelihood_extrinsic(true, &gp[i], method, lens.data(), fr.data(), da.data(), ps.data(),
                                              nullptr, "SIMPSONS", false, names.data(), ndet, 0.0);
    };
    /* untimed: the first call creates the resident event */
    out[0] = call(0);

    std::atomic<int> ready{0};
    std::atomic<bool> go{false};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&, t] {
            ready.fetch_add(1);
            while (!go.load(std::memory_order_acquire)) std::this_thread::yield();
            for (int r = 0; r < repeats; r++)
                for (int i = t; i < count; i += threads) out[i] = call(i);
        });
    while (ready.load() < threads) std::this_thread::yield();
    auto t0 = std::chrono::steady_clock::now();
    go.store(true, std::memory_order_release);
    for (auto &th : pool) th.join();
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    double st[9] = {0};
    (void)gwh_single_call_stats(st);
    if (handle >= 0) (void)gwh_host_data_release(handle);
    FILE *fo = fopen(argv[2], "wb");
    if (!fo || fwrite(out.data(), 8, (size_t)count, fo) != (size_t)count) return fail("cannot write the output");
    fclose(fo);
    const long long calls = (long long)count * repeats;
    printf("{\"threads\": %d, \"mode\": \"%s\", \"calls\": %lld, \"seconds\": %.6f, \"calls_per_s\": %.1f, "
           "\"batches\": %.0f, \"coalesced_instances\": %.0f, \"max_calls_per_batch\": %.0f}\n",
           threads, mode.c_str(), calls, secs, calls / secs, st[0], st[1], st[3]);
    return 0;
}
