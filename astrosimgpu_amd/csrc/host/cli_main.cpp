// astrosim-amd: the reference's command-line simulator (src/main.cpp) with the
// whole time step on the MI355X, through the C ABI.
//
// Same options, configuration files, outputs and file formats:
//   spikes.csv      time_ms,neuron          every recorded spike, in step
//                                           then source order
//   astrocytes.csv  time_ms,astrocyte,Ca,IP3  every record_every steps, the
//                                           first record_astro_max cells
//   neurons.csv     time_ms,neuron,V_m,I_SIC  likewise for neurons
//   run.txt         the run summary (format_summary, main.cpp:76-161) and the
//                   calcium analysis (main.cpp:264-313)
// Numbers are written with the stream's default format (6 significant
// digits) and times as step * dt, as recorder.cpp:41-58 and network.cpp:443
// do, so the files are byte-identical with the reference build's for the same
// configuration and seed (FP64).
//
// Differences, by design: the phase table of run.txt lists the device
// kernels' times instead of the host phases, and the calcium analysis covers
// the complete record (the reference analyses the file while its writer
// still holds the unflushed tail, so it can miss the last few kilobytes).
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "asg.h"

namespace {

void print_usage(const char* argv0) {
    std::cout
        << "Usage: " << argv0 << " [options]\n\n"
        << "Simulates a tripartite neuron-astrocyte network on the GPU and reports the\n"
        << "calcium and spiking statistics used to assess synchronisation.\n\n"
        << "Options:\n"
        << "  -c, --config PATH     parameter file (JSON)   [config/use_case.json]\n"
        << "  -o, --output DIR      output directory        [from config]\n"
        << "  -s, --seed N          random seed             [from config]\n"
        << "  -t, --sim-time MS     recorded duration       [from config]\n"
        << "  -p, --pre-time MS     discarded transient     [from config]\n"
        << "  -a, --astrocytes N    override the astrocyte count [from config]\n"
        << "      --fp32            FP32 state with FP64 islands (default FP64)\n"
        << "      --threads N       host threads for network construction\n"
        << "      --no-analysis     skip the post-run summary\n"
        << "  -h, --help            show this message\n";
}

struct Fail {
    std::string what;
};

void check(asg_status st) {
    if (st != ASG_OK) throw Fail{asg_last_error()};
}

}  // namespace

int main(int argc, char** argv) {
    std::string config_path = "config/use_case.json";
    std::string output_override;
    bool run_analysis = true;
    bool fp32 = false;
    asg_model_overrides ov{-1, -1.0, -1.0, -1, nullptr};

    try {
        for (int i = 1; i < argc; ++i) {
            const std::string arg = argv[i];
            auto next = [&](const char* what) -> std::string {
                if (i + 1 >= argc) throw Fail{std::string("missing value for ") + what};
                return argv[++i];
            };
            if (arg == "-h" || arg == "--help") {
                print_usage(argv[0]);
                return 0;
            } else if (arg == "-c" || arg == "--config") {
                config_path = next("--config");
            } else if (arg == "-o" || arg == "--output") {
                output_override = next("--output");
            } else if (arg == "-s" || arg == "--seed") {
                ov.seed = std::stoll(next("--seed"));
            } else if (arg == "-t" || arg == "--sim-time") {
                ov.sim_time = std::stod(next("--sim-time"));
            } else if (arg == "-p" || arg == "--pre-time") {
                ov.pre_time = std::stod(next("--pre-time"));
            } else if (arg == "-a" || arg == "--astrocytes") {
                ov.n_astro = std::stoll(next("--astrocytes"));
            } else if (arg == "--threads") {
                setenv("OMP_NUM_THREADS", next("--threads").c_str(), 1);
            } else if (arg == "--fp32") {
                fp32 = true;
            } else if (arg == "--no-analysis") {
                run_analysis = false;
            } else {
                std::cerr << "unknown option: " << arg << "\n\n";
                print_usage(argv[0]);
                return 2;
            }
        }

        std::cout << "Building network from " << config_path << " ...\n";
        asg_model* model = nullptr;
        check(asg_model_load(config_path.c_str(), &ov, &model));
        asg_model_info info{};
        check(asg_model_get_info(model, &info));
        const std::string out_dir = output_override.empty() ? info.output_dir : output_override;
        std::cout << "  " << info.n_astro << " astrocytes, " << info.n_exc << " excitatory, "
                  << info.n_inh << " inhibitory neurons\n"
                  << "  " << info.n_primary_exc + info.n_primary_inh << " neuron-neuron synapses, "
                  << info.n_neuron_to_astro << " neuron-astrocyte, " << info.n_astro_to_neuron
                  << " astrocyte-neuron\n";

        // Recorder(recorder.cpp:8-37): directory and headers.
        std::error_code ec;
        std::filesystem::create_directories(out_dir, ec);
        if (ec) throw Fail{"cannot create output directory " + out_dir + ": " + ec.message()};
        std::ofstream spikes_f, astro_f, neuron_f;
        if (info.record_spikes) {
            spikes_f.open(out_dir + "/spikes.csv");
            if (!spikes_f) throw Fail{"cannot open " + out_dir + "/spikes.csv"};
            spikes_f << "time_ms,neuron\n";
        }
        if (info.record_astro) {
            astro_f.open(out_dir + "/astrocytes.csv");
            if (!astro_f) throw Fail{"cannot open " + out_dir + "/astrocytes.csv"};
            astro_f << "time_ms,astrocyte,Ca,IP3\n";
        }
        if (info.record_neuron) {
            neuron_f.open(out_dir + "/neurons.csv");
            if (!neuron_f) throw Fail{"cannot open " + out_dir + "/neurons.csv"};
            neuron_f << "time_ms,neuron,V_m,I_SIC\n";
        }

        asg_net_desc desc = *asg_model_desc(model);
        desc.precision = fp32 ? ASG_FP32 : ASG_FP64;
        asg_net* net = nullptr;
        check(asg_net_create(&desc, &net));
        check(asg_net_set_profiling(net, 1));
        const int64_t pre = info.pre_steps, total = info.pre_steps + info.sim_steps;
        const int64_t every = info.record_every;
        const double dt = info.dt;
        const uint32_t n_neur = info.n_exc + info.n_inh;
        const uint32_t astro_rec = std::min(info.record_astro_max, info.n_astro);
        const uint32_t neuron_rec = std::min(info.record_neuron_max, n_neur);
        const int64_t ring = 4096;
        if (info.record_astro && astro_rec > 0) {
            check(asg_net_set_recording(net, pre, static_cast<int32_t>(every), astro_rec, ring));
        }

        std::cout << "Simulating " << info.pre_sim_time << " ms transient + " << info.sim_time
                  << " ms recorded ...\n";
        std::vector<int64_t> st(1 << 16);
        std::vector<uint32_t> src(1 << 16);
        std::vector<double> rec(static_cast<size_t>(ring) * 2 * std::max<uint32_t>(astro_rec, 1));
        std::vector<int64_t> rec_steps(static_cast<size_t>(ring));
        std::vector<double> V(neuron_rec), Isic(neuron_rec);
        std::vector<double> trace_t, trace_ca;  // astrocyte records kept for the analysis
        size_t spike_count = 0;

        auto drain = [&]() {
            int64_t got = 0;
            do {
                check(asg_net_spikes(net, st.data(), src.data(), static_cast<int64_t>(st.size()), &got));
                for (int64_t k = 0; k < got; ++k) {
                    if (st[k] < pre) continue;
                    if (info.record_spikes) {
                        spikes_f << static_cast<double>(st[k]) * dt << ',' << src[k] << '\n';
                        ++spike_count;
                    }
                }
            } while (got == static_cast<int64_t>(st.size()));
            if (info.record_astro && astro_rec > 0) {
                do {
                    check(asg_net_read_trace(net, rec.data(), rec_steps.data(), ring, &got));
                    for (int64_t r = 0; r < got; ++r) {
                        const double t_ms = static_cast<double>(rec_steps[r]) * dt;
                        const double* ca = rec.data() + static_cast<size_t>(r) * 2 * astro_rec;
                        for (uint32_t a = 0; a < astro_rec; ++a) {
                            astro_f << t_ms << ',' << a << ',' << ca[a] << ',' << ca[astro_rec + a] << '\n';
                        }
                        trace_t.push_back(t_ms);
                        trace_ca.insert(trace_ca.end(), ca, ca + astro_rec);
                    }
                } while (got == ring);
            }
        };

        const auto t0 = std::chrono::steady_clock::now();
        int64_t done = 0;
        while (done < total) {
            // Neuron records need the state after each recorded step, so
            // the device stops there; otherwise it runs in long chunks.
            int64_t stop = std::min(total, done + std::max<int64_t>(every, 1) * 256);
            if (info.record_neuron && neuron_rec > 0) {
                const int64_t next_rec = done <= pre ? pre : pre + ((done - pre + every - 1) / every) * every;
                stop = std::min(total, next_rec + 1);
            }
            check(asg_net_step(net, stop - done));
            done = stop;
            if (info.record_neuron && neuron_rec > 0) {
                const int64_t k = done - 1;
                if (k >= pre && (k - pre) % every == 0) {
                    check(asg_net_read_state(net, ASG_FIELD_V, 0, neuron_rec, V.data()));
                    check(asg_net_read_state(net, ASG_FIELD_I_SIC, 0, neuron_rec, Isic.data()));
                    const double t_ms = static_cast<double>(k) * dt;
                    for (uint32_t i = 0; i < neuron_rec; ++i) {
                        neuron_f << t_ms << ',' << i << ',' << V[i] << ',' << Isic[i] << '\n';
                    }
                }
            }
            drain();
        }
        const double wall =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        asg_kernel_times kt{};
        check(asg_net_kernel_times(net, &kt));
        spikes_f.close();
        astro_f.close();
        neuron_f.close();

        // Calcium analysis (main.cpp:264-313) over the complete record, each
        // value rounded to the 6 digits the file holds.
        std::string analysis;
        if (run_analysis && info.analysis_enabled && info.record_astro && astro_rec > 0) {
            const double threshold = info.ca_threshold >= 0.0 ? info.ca_threshold : info.SIC_th;
            asg_calcium_summary cs{};
            check(asg_analyze_calcium(trace_t.data(), trace_ca.data(), static_cast<int32_t>(astro_rec),
                                      static_cast<int64_t>(trace_t.size()), threshold,
                                      info.merge_gap, info.pre_sim_time, info.sim_time,
                                      info.hist_window, info.hist_shift, 1, &cs));
            std::ostringstream os;
            os << std::fixed << std::setprecision(4);
            os << "Calcium analysis (threshold " << threshold << " uM, window " << cs.width << "/"
               << cs.shift << " ms)\n";
            os << "-----------------------------------\n";
            os << "astrocytes with transients  " << cs.active << " / " << cs.cells << "\n";
            os << "transients detected         " << cs.transients << "\n";
            os << "mean duration               " << cs.mean_duration << " ms (sd " << cs.sd_duration
               << ")\n";
            os << "pairs correlated            " << cs.pairs << "\n";
            os << "mean pairwise correlation   " << cs.mean_correlation << "\n";
            analysis = os.str();
        }

        // format_summary, main.cpp:76-161
        std::ostringstream os;
        os << std::fixed << std::setprecision(4);
        os << "Tripartite neuron-astrocyte network\n";
        os << "-----------------------------------\n";
        os << "astrocytes            " << info.n_astro << "\n";
        os << "excitatory neurons    " << info.n_exc << "\n";
        os << "inhibitory neurons    " << info.n_inh << "\n";
        os << "exc -> neuron synapses " << info.n_primary_exc << "\n";
        os << "inh -> neuron synapses " << info.n_primary_inh << "\n";
        os << "neuron -> astrocyte    " << info.n_neuron_to_astro << "\n";
        os << "astrocyte -> neuron    " << info.n_astro_to_neuron << "\n";
        os << "\n";
        os << "dt                    " << dt << " ms (" << info.substeps << " substeps)\n";
        os << "astrocyte backend     MI355X whole step (" << (fp32 ? "FP32 + FP64 islands" : "FP64")
           << ")\n";
        os << "transient             " << info.pre_sim_time << " ms\n";
        os << "recorded              " << info.sim_time << " ms\n";
        os << "seed                  " << info.seed << "\n";
        os << "\n";
        os << "wall clock            " << wall << " s\n";
        os << "spikes recorded       " << spike_count << "\n";
        if (info.sim_time > 0.0 && n_neur > 0) {
            const double rate = static_cast<double>(spike_count) * 1000.0 /
                                (info.sim_time * static_cast<double>(n_neur));
            os << "mean firing rate      " << rate << " Hz\n";
        }
        const double dev = kt.ms[0] + kt.ms[1] + kt.ms[2] + kt.ms[3] + kt.ms[4];
        if (dev > 0.0) {
            os << "\n";
            os << "Device time by kernel (all steps)\n";
            os << "-----------------------------------\n";
            os << "                       seconds   % of device\n";
            auto row = [&](const char* name, double ms) {
                os << std::left << std::setw(22) << name << std::right << std::setw(8) << ms * 1e-3
                   << std::setw(11) << 100.0 * ms / dev << "\n";
                os << std::left;
            };
            row("external drive", kt.ms[4]);
            row("cell update", kt.ms[0]);
            row("spike compaction", kt.ms[1]);
            row("SIC + plasticity", kt.ms[2]);
            row("ordered delivery", kt.ms[3]);
            if (info.sim_time > 0.0) {
                os << "real-time factor      " << wall / ((info.pre_sim_time + info.sim_time) * 1e-3)
                   << "\n";
            }
        }
        if (!analysis.empty()) {
            os << "\n" << analysis;
        }
        const std::string summary = os.str();
        std::ofstream meta(out_dir + "/run.txt");
        if (meta) meta << summary;
        std::cout << "\n" << summary << "\nOutput written to " << out_dir << "/\n";
        asg_net_destroy(net);
        asg_model_free(model);
    } catch (const Fail& e) {
        std::cerr << "error: " << e.what << "\n";
        return 1;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
