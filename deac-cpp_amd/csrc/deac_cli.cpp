// deac.e -- command-line front end of the MI355X DEAC engine.
//
// Same flags, defaults, validation messages, exit codes, output files and
// log keys as the reference's main() (deac.cpp:1925-2197) and its output
// section (deac.cpp:1687-1769).  The reference's compile-time variants are
// runtime flags here: --model (DEAC_MODEL), --zero_temperature (the ZeroT
// build, files prefixed deac-zT), --spectra_type bdsf (the bosonic
// detailed-balance build, with --use_negative_first_moment),
// --allow_negative_spectral_weight and
// --single_particle_fermionic_spectral_function.
//
// Seed batches: --number_of_seeds K runs seeds seed .. seed+K-1 in one
// batched launch (deac_batch_run) and writes, for each seed, exactly the
// files and log of a separate `deac.e --seed s` run -- the workflow of
// tools/generate_commands.sh:18-32 in one process.  --average_dsf adds the
// seed-averaged DSF, summed in the order tools/plot_simple_average_deac_results.py:8-28
// uses (run keys sorted as strings).
//
// --ks_precision fp32 | bf16x3 runs the reconstruction with the model ISF's
// K.S contraction in that precision; the random draws, chi^2, selection and
// statistics stay the reference's, in fp64.  One seed runs the split
// generation loop; a seed batch runs the batch kernel's delta form (E of the
// parent plus K times the trial's change, FP64 re-evaluation every 1024
// generations and at the end), so at these precisions a batched seed follows
// a trajectory of its own, not that of a single --seed run.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "deac_engine.h"

namespace fs = std::filesystem;

namespace {

[[noreturn]] void fail_with_error(const std::string &msg)  // deac.cpp:35-38
{
    std::cerr << msg << std::endl;
    std::exit(1);
}

// Raw float64 file (deac.cpp:58-100): exit 1 on open/size errors, 2 on
// allocation failure, 3 on a short read.
std::vector<double> load_doubles(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "File error: %s\n", path.c_str());
        std::exit(1);
    }
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::rewind(f);
    if (bytes < 0 || bytes % static_cast<long>(sizeof(double)) != 0) {
        std::fclose(f);
        fail_with_error("File error: " + path + " does not contain a whole number of doubles");
    }
    const size_t n = static_cast<size_t>(bytes) / sizeof(double);
    if (n == 0) {
        std::fclose(f);
        fail_with_error("File error: " + path + " is empty");
    }
    std::vector<double> v;
    try {
        v.resize(n);
    } catch (const std::bad_alloc &) {
        std::fputs("Memory error", stderr);
        std::exit(2);
    }
    if (std::fread(v.data(), 1, static_cast<size_t>(bytes), f) != static_cast<size_t>(bytes)) {
        std::fputs("Reading error", stderr);
        std::exit(3);
    }
    std::fclose(f);
    return v;
}

void write_array(const fs::path &path, const double *data, size_t n)  // deac.cpp:51-56
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) fail_with_error("File error: cannot write " + path.string());
    std::fwrite(data, sizeof(double), n, f);
    std::fclose(f);
}

// ---------------------------------------------------------------- arguments

struct Option {
    std::string short_name, long_name, help, def;  // def: text shown as [default: ...]
    bool flag = false;                               // implicit-true switch
    bool present = false;
    std::string value;
};

class Parser {
public:
    std::vector<Option> opts;
    std::string positional_value;
    bool positional_present = false;

    Option &add(std::string s, std::string l, std::string help, std::string def, bool flag = false)
    {
        opts.push_back(Option{std::move(s), std::move(l), std::move(help), std::move(def), flag, false, ""});
        return opts.back();
    }
    Option *find(const std::string &name)
    {
        // A later registration of the same short name wins, as in the
        // reference (-v selects the differential-weight probability).
        for (auto it = opts.rbegin(); it != opts.rend(); ++it)
            if (it->short_name == name || it->long_name == name) return &*it;
        return nullptr;
    }
    const Option &get(const std::string &long_name) const
    {
        for (const auto &o : opts)
            if (o.long_name == long_name) return o;
        throw std::logic_error("no option " + long_name);
    }
    std::string usage() const
    {
        std::ostringstream u;
        u << "Usage: deac-cpp [--help] [--version]";
        for (const auto &o : opts) u << " [" << o.long_name << (o.flag ? "" : " VAR") << "]";
        u << " isf_file\n\nPositional arguments:\n";
        auto row = [&u](const std::string &names, const std::string &help) {
            u << "  " << names;
            for (size_t k = names.size(); k < 51; k++) u << ' ';
            u << "\t" << help << " \n";
        };
        row("isf_file", "binary file containing isf data (tau, isf, error)");
        u << "\nOptional arguments:\n";
        row("-h, --help", "shows help message and exits");
        row("--version", "prints version information and exits");
        for (const auto &o : opts) {
            const std::string names = o.short_name.empty() ? o.long_name : o.short_name + ", " + o.long_name;
            row(names, o.def.empty() ? o.help : o.help + " [default: " + o.def + "]");
        }
        return u.str();
    }
    void parse(int argc, char **argv)
    {
        for (int k = 1; k < argc; k++) {
            std::string a = argv[k];
            if (a == "-h" || a == "--help") {
                std::cout << usage() << std::endl;
                std::exit(0);
            }
            if (a == "--version") {
                std::cout << DEAC_ENGINE_VERSION << std::endl;
                std::exit(0);
            }
            if (a.size() > 1 && a[0] == '-' && !(std::isdigit(static_cast<unsigned char>(a[1])) || a[1] == '.')) {
                std::string name = a, val;
                const bool has_eq = a.rfind("--", 0) == 0 && a.find('=') != std::string::npos;
                if (has_eq) {
                    name = a.substr(0, a.find('='));
                    val = a.substr(a.find('=') + 1);
                }
                Option *o = find(name);
                if (!o) throw std::runtime_error("Unknown argument: " + a);
                o->present = true;
                if (o->flag) {
                    o->value = "1";
                    continue;
                }
                if (!has_eq) {
                    if (k + 1 >= argc) throw std::runtime_error("Too few arguments for '" + o->long_name + "'.");
                    val = argv[++k];
                }
                o->value = val;
            } else {
                if (positional_present) throw std::runtime_error("Maximum number of positional arguments exceeded");
                positional_present = true;
                positional_value = a;
            }
        }
        if (!positional_present) throw std::runtime_error("isf_file: 1 argument(s) expected. 0 provided.");
    }
};

double to_double(const Option &o, double def)
{
    if (!o.present) return def;
    try {
        size_t pos = 0;
        const double v = std::stod(o.value, &pos);
        return v;
    } catch (const std::exception &) {
        throw std::runtime_error("Invalid value for " + o.long_name + ": " + o.value);
    }
}

unsigned long to_ulong(const Option &o, unsigned long def)
{
    if (!o.present) return def;
    try {
        return std::stoul(o.value);
    } catch (const std::exception &) {
        throw std::runtime_error("Invalid value for " + o.long_name + ": " + o.value);
    }
}

const char *model_build_name(int model)
{
    return model == DEAC_MODEL_HYPERBOLIC ? "USE_HYPERBOLIC_MODEL"
           : model == DEAC_MODEL_NORMALIZATION ? "USE_NORMALIZATION_MODEL"
                                               : "USE_STANDARD_MODEL";
}

}  // namespace

int main(int argc, char **argv)
{
    Parser prog;
    prog.add("-T", "--temperature", "Temperature of system.", "0");
    prog.add("-N", "--number_of_generations", "Number of generations before genetic algorithm quits.", "100000");
    prog.add("-P", "--population_size", "Size of initial population", "512");
    prog.add("-M", "--genome_size", "Size of genome.", "512");
    prog.add("", "--omega_max", "Maximum frequency to explore.", "60");
    prog.add("", "--frequency_file", "Filename containing frequency partition (genome_size and omega_max will be ignored).", "");
    prog.add("", "--normalize", "Normalize spectrum to the zeroeth moment.", "", true);
    prog.add("", "--use_negative_first_moment", "Calculate inverse first moment from ISF data and use it in fitness.", "", true);
    prog.add("", "--first_moment", "Set first frequency moment and use in fitness function.", "-1");
    prog.add("", "--third_moment", "Set third frequency moment and use in fitness function.", "-1");
    prog.add("", "--third_moment_error", "Set error for third frequency moment.", "0");
    prog.add("-r", "--crossover_probability", "Initial probability for parent gene to become mutant vector gene.", "0.9");
    prog.add("-u", "--self_adapting_crossover_probability", "Probability for `crossover_probability` to mutate.", "0.1");
    prog.add("-F", "--differential_weight", "Initial weight factor when creating mutant vector.", "0.9");
    prog.add("-v", "--self_adapting_differential_weight_probability", "Probability for `differential_weight` to mutate.", "0.1");
    prog.add("-l", "--self_adapting_differential_weight_shift", "If `self_adapting_differential_weight_probability` mutate, new value is `l + m*rand()`.", "0.1");
    prog.add("-m", "--self_adapting_differential_weight", "If `self_adapting_differential_weight_probability` mutate, new value is `l + m*rand()`.", "0.9");
    prog.add("", "--stop_minimum_fitness", "Stopping criteria, if minimum fitness is below `stop_minimum_fitness` stop evolving.", "1");
    prog.add("", "--seed", "Seed to pass to random number generator.", "0");
    prog.add("", "--uuid", "UUID for run. If empty will be set to `seed`.", "\"\"");
    prog.add("", "--save_state", "Save state of DEAC algorithm. Saves the random number generator, population, and population fitness.", "", true);
    prog.add("", "--save_directory", "Directory to save results in.", "\"./deacresults\"");
    prog.add("", "--track_stats", "Track minimum fitness and other stats.", "", true);
    prog.add("", "--spectra_type", "Choose spectral type for kernel factors [spbsf, spfsf, bfull, ffull, bdsf].", "\"spfsf\"");
    prog.add("", "--model", "Spectral model [standard, hyperbolic, normalization].", "\"standard\"");
    prog.add("", "--zero_temperature", "Zero-temperature kernel exp(-tau*omega), positive frequencies only.", "", true);
    prog.add("", "--allow_negative_spectral_weight", "Allow negative spectral weights.", "", true);
    prog.add("", "--single_particle_fermionic_spectral_function", "Renormalize every generation when --normalize is set.", "", true);
    prog.add("", "--number_of_seeds", "Run seeds seed, seed+1, .., seed+K-1 in one batched launch; each seed writes the files of its own run.", "1");
    prog.add("", "--average_dsf", "With --number_of_seeds: also write the seed-averaged DSF (<prefix>_seed-average_<first>-<last>.bin).", "", true);
    prog.add("", "--ks_precision", "Precision of the model ISF K.S [fp64, fp32, bf16x3]; chi^2 and selection stay fp64.", "\"fp64\"");

    try {
        prog.parse(argc, argv);
    } catch (const std::runtime_error &err) {  // deac.cpp:2025-2032
        std::cout << err.what() << std::endl;
        std::cout << prog.usage() << std::endl;
        std::exit(1);
    }

    try {
        const unsigned long seed_arg = to_ulong(prog.get("--seed"), 0);
        const unsigned long nseeds = to_ulong(prog.get("--number_of_seeds"), 1);
        if (nseeds < 1) fail_with_error("number_of_seeds must be at least 1");
        std::string uuid = prog.get("--uuid").present ? prog.get("--uuid").value : "";
        if (nseeds > 1 && !uuid.empty()) fail_with_error("--uuid cannot be combined with --number_of_seeds > 1");
        if (uuid.empty()) uuid = std::to_string(seed_arg);  // deac.cpp:2034-2037
        if (nseeds == 1) std::cout << "uuid: " << uuid << std::endl;

        const std::string spectra = prog.get("--spectra_type").present ? prog.get("--spectra_type").value : "spfsf";
        // bdsf: the reference's bosonic detailed-balance build (deac.cpp:2016-2018)
        const std::map<std::string, int> spectra_ids = {{"spbsf", DEAC_SPBSF}, {"spfsf", DEAC_SPFSF}, {"bfull", DEAC_BFULL},
                                                        {"ffull", DEAC_FFULL}, {"bdsf", DEAC_BDSF}};
        if (!spectra_ids.count(spectra)) {  // deac.cpp:2040-2058
            std::cout << "Please choose spectra_type from the following options: spbsf, spfsf, bfull, ffull, bdsf" << std::endl;
            std::exit(1);
        }
        const std::string model_name = prog.get("--model").present ? prog.get("--model").value : "standard";
        const std::map<std::string, int> model_ids = {{"standard", DEAC_MODEL_STANDARD}, {"hyperbolic", DEAC_MODEL_HYPERBOLIC}, {"normalization", DEAC_MODEL_NORMALIZATION}};
        if (!model_ids.count(model_name)) {
            std::cout << "Please choose model from the following options: standard, hyperbolic, normalization" << std::endl;
            std::exit(1);
        }
        const bool zero_t = prog.get("--zero_temperature").present;

        const std::string isf_file = prog.positional_value;
        std::vector<double> data = load_doubles(isf_file);  // deac.cpp:2060-2074
        if (data.size() % 3 != 0) fail_with_error("ISF input file must contain tau, isf, and error arrays of equal length");
        const size_t T = data.size() / 3;
        if (T < 2) fail_with_error("ISF input file must contain at least two timeslices");
        const double *tau = data.data(), *isf = data.data() + T, *err = data.data() + 2 * T;
        for (size_t i = 0; i < T; i++) {
            if (!std::isfinite(tau[i]) || !std::isfinite(isf[i]) || !std::isfinite(err[i]))
                fail_with_error("ISF input file contains non-finite values");
            if ((spectra == "spbsf" || spectra == "spfsf") && isf[i] > 0.0)
                fail_with_error("positive ISF values are not supported for single-particle spectra");
            if (err[i] <= 0.0) fail_with_error("ISF input errors must be positive");
        }


        double temperature = to_double(prog.get("--temperature"), 0.0);
        if (!zero_t && (!std::isfinite(temperature) || temperature <= 0.0))
            fail_with_error("temperature must be positive for non-ZeroT builds");
        const size_t N = to_ulong(prog.get("--number_of_generations"), 100000);
        const size_t P = to_ulong(prog.get("--population_size"), 512);
        if (N < 2) fail_with_error("number_of_generations must be at least 2");
        if (P < 4) fail_with_error("population_size must be at least 4");

        std::vector<double> frequency;
        if (prog.get("--frequency_file").present) {
            frequency = load_doubles(prog.get("--frequency_file").value);
        } else {
            const size_t G = to_ulong(prog.get("--genome_size"), 512);
            const double omega_max = to_double(prog.get("--omega_max"), 60.0);
            if (G < 2) fail_with_error("genome_size must be at least 2");
            if (!std::isfinite(omega_max) || omega_max <= 0.0) fail_with_error("omega_max must be positive");
            frequency.resize(G);
            deac_frequency_grid(G, omega_max, frequency.data());
        }
        if (frequency.size() < 2) fail_with_error("frequency_file must contain at least two frequencies");
        for (size_t i = 0; i < frequency.size(); i++) {
            if (!std::isfinite(frequency[i]) || frequency[i] < 0.0) fail_with_error("frequencies must be finite and non-negative");
            if (i > 0 && frequency[i] < frequency[i - 1]) fail_with_error("frequencies must be sorted in non-decreasing order");
        }

        const bool normalize = prog.get("--normalize").present;
        bool use_negative_first_moment = prog.get("--use_negative_first_moment").present;
        const double first_moment = to_double(prog.get("--first_moment"), -1.0);
        const double third_moment = to_double(prog.get("--third_moment"), -1.0);
        const double third_moment_error = to_double(prog.get("--third_moment_error"), 0.0);
        if (!std::isfinite(first_moment)) fail_with_error("first_moment must be finite");
        if (!std::isfinite(third_moment)) fail_with_error("third_moment must be finite");
        if (third_moment >= 0.0 && (!std::isfinite(third_moment_error) || third_moment_error <= 0.0))
            fail_with_error("third_moment_error must be positive when third_moment is used");

        deac_problem p{};
        p.number_of_timeslices = T;
        p.imaginary_time = tau;
        p.isf = isf;
        p.isf_error = err;
        p.genome_size = frequency.size();
        p.frequency = frequency.data();
        p.temperature = temperature;
        p.model = model_ids.at(model_name);
        p.spectra_type = spectra_ids.at(spectra);
        p.zero_temperature = zero_t;
        p.allow_negative_spectral_weight = prog.get("--allow_negative_spectral_weight").present;
        p.renormalize_each_generation = prog.get("--single_particle_fermionic_spectral_function").present;
        p.population_size = P;
        p.normalize = normalize;
        p.first_moment = first_moment;
        p.third_moment = third_moment;
        p.third_moment_error = third_moment_error;
        p.crossover_probability = to_double(prog.get("--crossover_probability"), 0.9);
        p.self_adapting_crossover_probability = to_double(prog.get("--self_adapting_crossover_probability"), 0.1);
        p.differential_weight = to_double(prog.get("--differential_weight"), 0.9);
        p.self_adapting_differential_weight_probability = to_double(prog.get("--self_adapting_differential_weight_probability"), 0.1);
        p.use_negative_first_moment = use_negative_first_moment && !zero_t;  // deac.cpp:230-236
        const double shift = to_double(prog.get("--self_adapting_differential_weight_shift"), 0.1);
        const double sadw = to_double(prog.get("--self_adapting_differential_weight"), 0.9);
        const double stop = to_double(prog.get("--stop_minimum_fitness"), 1.0);
        const bool track_stats = prog.get("--track_stats").present;

        const fs::path save_directory(prog.get("--save_directory").present ? prog.get("--save_directory").value : "./deacresults");
        std::error_code ec;
        fs::create_directory(save_directory, ec);  // deac.cpp:2167
        if (ec) fail_with_error("File error: cannot create save directory " + save_directory.string() + ": " + ec.message());

        const std::string prefix = std::string(zero_t ? "deac-zT" : "deac-") + spectra;  // deac.cpp:1687-1691
        std::vector<std::string> uuids(nseeds);
        for (unsigned long k = 0; k < nseeds; k++) uuids[k] = nseeds == 1 ? uuid : std::to_string(seed_arg + k);
        auto log_path_of = [&](const std::string &id) { return save_directory / (prefix + "_log_" + id + ".dat"); };
        for (unsigned long k = 0; k < nseeds; k++) {
            std::ofstream log(log_path_of(uuids[k]), std::ios_base::out | std::ios_base::app);  // deac.cpp:2169-2182
            log << "uuid: " << uuids[k] << std::endl;
            log << "isf_file: " << isf_file << std::endl;
        }

        if (zero_t) {  // deac.cpp:230-236
            if (use_negative_first_moment) std::cout << "use_negative_first_moment disabled for zero temperature build" << std::endl;
            use_negative_first_moment = false;
            temperature = 0.0;
        }
        if (first_moment >= 0.0) std::cout << "WARNING: setting first_moment_error to 1.0." << std::endl;  // deac.cpp:627-628

        // One result per seed: the batch kernel when the problem fits it,
        // else the single-instance engine seed by seed (same results).
        const size_t nout = deac_output_length(&p);
        std::vector<std::vector<double>> dsf(nseeds, std::vector<double>(nout)), freq_out(nseeds, std::vector<double>(nout));
        std::vector<std::vector<double>> s_mean(nseeds), s_min(nseeds), s_sq(nseeds);
        std::vector<deac_result> res(nseeds);
        std::vector<uint64_t> states(4 * nseeds);
        for (unsigned long k = 0; k < nseeds; k++) {
            deac_seed_state(seed_arg + k, &states[4 * k]);  // deac.cpp:2089-2091
            res[k] = deac_result{};
            res[k].dsf = dsf[k].data();
            res[k].frequency = freq_out[k].data();
            if (track_stats) {
                s_mean[k].resize(N);
                s_min[k].resize(N);
                s_sq[k].resize(N);
                res[k].stats_mean = s_mean[k].data();
                res[k].stats_minimum = s_min[k].data();
                res[k].stats_squared_mean = s_sq[k].data();
            }
        }
        const std::string ksp = prog.get("--ks_precision").present ? prog.get("--ks_precision").value : "fp64";
        const std::map<std::string, int> ks_ids = {{"fp64", DEAC_PRECISION_FP64}, {"fp32", DEAC_PRECISION_FP32},
                                                   {"bf16x3", DEAC_PRECISION_BF16X3}};
        if (!ks_ids.count(ksp)) {
            std::cout << "Please choose ks_precision from the following options: fp64, fp32, bf16x3" << std::endl;
            std::exit(1);
        }
        const deac_run_options ro{N, stop, track_stats ? 1 : 0, ks_ids.at(ksp), 0};
        const char *msg = nullptr;
        bool done = false;
        if (nseeds > 1) {
            std::vector<deac_problem> probs(nseeds, p);
            deac_batch *b = nullptr;
            deac_status st = deac_batch_create(probs.data(), nseeds, &b, &msg);
            if (st == DEAC_OK) {
                std::vector<deac_run_options> opts(nseeds, ro);
                st = deac_batch_run(b, states.data(), opts.data(), res.data());
                deac_batch_destroy(b);
                if (st == DEAC_ERR_UNSUPPORTED) {  // refused before any launch (e.g. mixed K.S with moments)
                    msg = deac_last_error_message();
                } else if (st != DEAC_OK) {
                    std::cerr << "deac engine: " << deac_status_string(st) << std::endl;
                    std::exit(1);
                } else {
                    done = true;
                }
            } else if (st != DEAC_ERR_UNSUPPORTED) {
                std::cerr << "deac engine: " << deac_status_string(st) << (msg ? std::string(": ") + msg : "") << std::endl;
                std::exit(1);
            }
            if (!done) std::cout << "seeds run one after another (" << (msg ? msg : "not batchable") << ")" << std::endl;
        }
        if (!done) {
            deac_handle *h = nullptr;
            deac_status st = deac_create(&p, &h, &msg);
            if (st != DEAC_OK) {
                std::cerr << "deac engine: " << deac_status_string(st) << (msg ? std::string(": ") + msg : "") << std::endl;
                std::exit(1);
            }
            for (unsigned long k = 0; k < nseeds && st == DEAC_OK; k++) st = deac_run(h, &states[4 * k], &ro, &res[k]);
            deac_destroy(h);
            if (st != DEAC_OK) {
                std::cerr << "deac engine: " << deac_status_string(st) << std::endl;
                std::exit(1);
            }
        }

        // Output files and log of each seed (deac.cpp:1687-1769).
        for (unsigned long k = 0; k < nseeds; k++) {
            const std::string &id = uuids[k];
            const deac_result &r = res[k];
            if (nseeds > 1) std::cout << "uuid: " << id << std::endl;
            const fs::path dsf_path = save_directory / (prefix + "_dsf_" + id + ".bin");
            const fs::path freq_path = save_directory / (prefix + "_frequency_" + id + ".bin");
            write_array(dsf_path, dsf[k].data(), nout);
            write_array(freq_path, freq_out[k].data(), nout);
            fs::path mean_path, min_path, sq_path;
            if (track_stats) {
                mean_path = save_directory / (prefix + "_stats_fitness-mean_" + id + ".bin");
                min_path = save_directory / (prefix + "_stats_fitness-minimum_" + id + ".bin");
                sq_path = save_directory / (prefix + "_stats_fitness-squared-mean_" + id + ".bin");
                write_array(mean_path, s_mean[k].data(), r.generation + 1);
                write_array(min_path, s_min[k].data(), r.generation + 1);
                write_array(sq_path, s_sq[k].data(), r.generation + 1);
            }
            std::ofstream log(log_path_of(id), std::ios_base::out | std::ios_base::app);
            log << "build: " << model_build_name(p.model) << std::endl;
            log << "kernel: " << spectra << std::endl;
            log << "temperature: " << temperature << std::endl;
            log << "number_of_generations: " << N << std::endl;
            log << "number_of_timeslices: " << T << std::endl;
            log << "population_size: " << P << std::endl;
            log << "genome_size: " << frequency.size() << std::endl;
            log << "normalize: " << normalize << std::endl;
            log << "use_negative_first_moment: " << use_negative_first_moment << std::endl;
            log << "first_moment: " << first_moment << std::endl;
            log << "third_moment: " << third_moment << std::endl;
            log << "third_moment_error: " << third_moment_error << std::endl;
            log << "crossover_probability: " << p.crossover_probability << std::endl;
            log << "self_adapting_crossover_probability: " << p.self_adapting_crossover_probability << std::endl;
            log << "differential_weight: " << p.differential_weight << std::endl;
            log << "self_adapting_differential_weight_probability: " << p.self_adapting_differential_weight_probability << std::endl;
            log << "self_adapting_differential_weight_shift: " << shift << std::endl;
            log << "self_adapting_differential_weight: " << sadw << std::endl;
            log << "stop_minimum_fitness: " << stop << std::endl;
            log << "track_stats: " << track_stats << std::endl;
            log << "seed: " << seed_arg + k << std::endl;
            log << "best_dsf_filename: " << dsf_path << std::endl;
            log << "frequency_filename: " << freq_path << std::endl;
            log << "generation: " << r.generation << std::endl;
            std::cout << "generation: " << r.generation << std::endl;
            log << "minimum_fitness: " << r.minimum_fitness << std::endl;
            std::cout << "minimum_fitness: " << r.minimum_fitness << std::endl;
            if (track_stats) {
                log << "fitness_mean_filename: " << mean_path << std::endl;
                log << "fitness_minimum_filename: " << min_path << std::endl;
                log << "fitness_squared_mean_filename: " << sq_path << std::endl;
            }
        }

        // Seed average (tools/plot_simple_average_deac_results.py:8-28): the
        // first key's DSF plus the others in string order of the keys, then
        // divided by the count.
        if (prog.get("--average_dsf").present) {
            std::vector<unsigned long> order(nseeds);
            for (unsigned long k = 0; k < nseeds; k++) order[k] = k;
            std::sort(order.begin(), order.end(), [&](unsigned long a, unsigned long b) { return uuids[a] < uuids[b]; });
            std::vector<double> avg(dsf[order[0]]);
            for (unsigned long k = 1; k < nseeds; k++)
                for (size_t i = 0; i < nout; i++) avg[i] += dsf[order[k]][i];
            for (size_t i = 0; i < nout; i++) avg[i] /= static_cast<double>(nseeds);
            const fs::path avg_path = save_directory / (prefix + "_seed-average_" + std::to_string(seed_arg) + "-" +
                                                        std::to_string(seed_arg + nseeds - 1) + ".bin");
            write_array(avg_path, avg.data(), nout);
            std::cout << "seed_average_filename: " << avg_path << std::endl;
        }
    } catch (const std::runtime_error &err) {
        std::cout << err.what() << std::endl;
        std::cout << prog.usage() << std::endl;
        std::exit(1);
    }
    return 0;
}
