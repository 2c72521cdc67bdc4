// sdca_cli.cpp -- SuperDCA-compatible command line on top of libsdca.so.
//
// Mirrors SuperDCA's main pipeline (SuperDCA/src/SuperDCA.cpp:40-421) and option set
// (SuperDCA/src/SuperDCA_options.cpp:159-189, SuperDCA/src/plmDCA_options.cpp:135-152,
// apegrunt's --maf-threshold/--gap-threshold/--no-filter-alignment of README.md:32):
// FASTA -> MAF/gap filter -> envelope states -> weights -> per-locus solve on the GPU
// -> "<id>.SuperDCA_couplings.<base>-based.all" (plmDCA.hpp:589-697).  GPU options form
// their own group (--device, --batch-loci, --precision, --history, --q-env).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "sdca.h"

namespace {

struct Options {
    bool verbose = false;
    std::vector<std::string> alignments;
    std::string filterlist, samplelist, locilist;
    bool out_filtered = false, out_filterlist = false, out_samplelist = false;
    int threads = -1;
    bool no_reweighting = false;
    double reweighting_threshold = 0.9;  // plmDCA_options.cpp:47
    bool output_weights = false;
    bool norm_of_mean = false;
    double gradient_threshold = 1e-3;  // plmDCA_options.cpp:61
    double coupling_threshold = 0.0;
    double lambda_h = -1.0, lambda_J = -1.0;
    bool no_dca = false, no_coupling_output = false;
    long output_top = 0;                // > 0: also write the largest scores (.top)
    double output_min_score = -HUGE_VAL;  // and/or only scores at least this large
    double maf_threshold = 0.01, gap_threshold = 0.15;  // README.md:32
    bool no_filter = false;
    int output_base = 1, input_base = 1;
    long begin_locus = 0, end_locus = -1;
    int device = 0, batch_loci = 0, history = 0, q_env = 4, max_iter = 2000;
    int precision = SDCA_PREC_FP32;
    int formulation = SDCA_FORM_AUTO;
    int chunk_loci = 0;  // loci per solve call (0: 4096, or fewer with --norm-of-mean-scoring)
    int subspace_rank = 128;  // Gram: warm start from the dominant eigen-subspace (0: cold)
};

const char *kUsage =
    "Usage: sdca [options] <alignmentfile> [<second alignmentfile>]\n"
    "  (with two alignments: scan every locus of the first against the loci of the second)\n"
    "General:\n"
    "  -h [ --help ]                  Print this help message.\n"
    "  -v [ --verbose ]               Be verbose.\n"
    "  --filterlistfile arg           Loci filter list (columns of the input alignment to keep).\n"
    "  --output-filtered-alignment    Write the filtered alignment (.fasta) and its loci list (.loci).\n"
    "  --output-filterlist-alignment  Write the alignment after filter-list selection.\n"
    "  --samplelistfile arg           Sample list (sequence names to keep).\n"
    "  --output-samplelist-alignment  Write the alignment after sample-list selection.\n"
    "  --locilistfile arg             Loci (columns of the filtered alignment) to analyse.\n"
    "  -t [ --threads ] arg           Host threads (accepted for compatibility; the solve runs on the GPU).\n"
    "Alignment:\n"
    "  --maf-threshold arg (=0.01)    Minimum minor-allele frequency.\n"
    "  --gap-threshold arg (=0.15)    Maximum gap frequency.\n"
    "  --no-filter-alignment          Do not filter loci.\n"
    "  --output-indexing-base arg (=1)\n"
    "  --input-indexing-base arg (=1)\n"
    "  --begin-locus arg (=0)  --end-locus arg (=-1)\n"
    "plmDCA:\n"
    "  --no-reweighting               Do not reweight samples.\n"
    "  --reweighting-threshold arg (=0.9)\n"
    "  --output-weights               Write sample weights to file.\n"
    "  --norm-of-mean-scoring         Score the mean of the two gauged couplings (six columns; stores\n"
    "                                 every q x q block on disk: loci^2 x q^2 x 4 bytes).\n"
    "  --gradient-threshold arg (=0.001)\n"
    "  --coupling-gradient-threshold arg (=0)\n"
    "                                 Also require the couplings' gradient below arg (0: off):\n"
    "                                 ||g_J||_inf (primal) or ||g_J||_2 (gram); 3e-6 or 1e-5\n"
    "                                 puts the scores within 1e-4 of the exact optimum.\n"
    "  --lambda-h arg (=-1)           (< 0: automatic)\n"
    "  --lambda-J arg (=-1)           (< 0: automatic)\n"
    "  --no-dca                       Stop after the weights.\n"
    "  --no-coupling-output           Do not write coupling scores.\n"
    "  --output-top arg               Also write the arg largest scores, largest first (.top).\n"
    "  --output-min-score arg         Also write every score >= arg, largest first (.top).\n"
    "GPU:\n"
    "  --device arg (=0)  --history arg (=0: 32 Gram, 10 primal)  --max-iter arg (=2000)\n"
    "  --batch-loci arg (=0)          Loci in flight (0: 3072 for the Gram formulation, 128 otherwise).\n"
    "  --formulation auto|gram|primal (=auto)\n"
    "                                 gram: couplings in the span of the sequences, FP64, one B x B\n"
    "                                 GEMM per evaluation, stopping on 2-norms; primal: the q x K_x\n"
    "                                 parameters (--precision); auto: gram when B <= K_x.\n"
    "  --subspace-rank arg (=128)     Gram: start each locus from its solve in the dominant 128-dim\n"
    "                                 eigen-subspace of the identical-locus counts (0: from zero).\n"
    "  --chunk-loci arg (=4096)       Loci per solve call; the scores stream through a file next to\n"
    "                                 the output, so host memory is O(chunk x loci).\n"
    "  --precision fp32|split-bf16 (=fp32)  --q-env arg (=4; 5 keeps four alleles plus gap)\n";

[[noreturn]] void die(const std::string &msg)
{
    std::cerr << "SuperDCA error: " << msg << "\n\n";
    std::exit(EXIT_FAILURE);
}

bool file_exists(const std::string &p)
{
    struct stat sb;
    return stat(p.c_str(), &sb) == 0;
}

// get_unique_ofstream (SuperDCA_commons.h:65-81): append .1, .2, ... if the name exists.
std::string unique_name(const std::string &base)
{
    if (!file_exists(base)) return base;
    for (int k = 1;; ++k) {
        const std::string cand = base + "." + std::to_string(k);
        if (!file_exists(cand)) return cand;
    }
}

std::string id_string(const std::string &path)
{
    std::string b = path.substr(path.find_last_of('/') + 1);
    const auto dot = b.find_last_of('.');
    if (dot != std::string::npos && dot > 0) b = b.substr(0, dot);
    return b;
}

Options parse_args(int argc, char **argv)
{
    Options o;
    auto need = [&](int &i) -> std::string {
        if (i + 1 >= argc) die(std::string("option ") + argv[i] + " needs a value");
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::cout << "sdca: plmDCA on AMD MI355X (SuperDCA-compatible command line)\n" << kUsage;
            std::exit(EXIT_SUCCESS);
        } else if (a == "-v" || a == "--verbose") o.verbose = true;
        else if (a == "--alignmentfile") o.alignments.push_back(need(i));
        else if (a == "--filterlistfile") o.filterlist = need(i);
        else if (a == "--output-filtered-alignment") o.out_filtered = true;
        else if (a == "--output-filterlist-alignment") o.out_filterlist = true;
        else if (a == "--samplelistfile") o.samplelist = need(i);
        else if (a == "--output-samplelist-alignment") o.out_samplelist = true;
        else if (a == "--locilistfile") o.locilist = need(i);
        else if (a == "-t" || a == "--threads") o.threads = std::atoi(need(i).c_str());
        else if (a == "--no-reweighting") o.no_reweighting = true;
        else if (a == "--reweighting-threshold") o.reweighting_threshold = std::atof(need(i).c_str());
        else if (a == "--output-weights") o.output_weights = true;
        else if (a == "--norm-of-mean-scoring") o.norm_of_mean = true;
        else if (a == "--gradient-threshold") o.gradient_threshold = std::atof(need(i).c_str());
        else if (a == "--coupling-gradient-threshold") o.coupling_threshold = std::atof(need(i).c_str());
        else if (a == "--lambda-h") o.lambda_h = std::atof(need(i).c_str());
        else if (a == "--lambda-J") o.lambda_J = std::atof(need(i).c_str());
        else if (a == "--no-dca") o.no_dca = true;
        else if (a == "--no-coupling-output") o.no_coupling_output = true;
        else if (a == "--output-top") o.output_top = std::atol(need(i).c_str());
        else if (a == "--output-min-score") o.output_min_score = std::atof(need(i).c_str());
        else if (a == "--maf-threshold") o.maf_threshold = std::atof(need(i).c_str());
        else if (a == "--gap-threshold") o.gap_threshold = std::atof(need(i).c_str());
        else if (a == "--no-filter-alignment") o.no_filter = true;
        else if (a == "--output-indexing-base") o.output_base = std::atoi(need(i).c_str());
        else if (a == "--input-indexing-base") o.input_base = std::atoi(need(i).c_str());
        else if (a == "--begin-locus") o.begin_locus = std::atol(need(i).c_str());
        else if (a == "--end-locus") o.end_locus = std::atol(need(i).c_str());
        else if (a == "--device") o.device = std::atoi(need(i).c_str());
        else if (a == "--batch-loci") o.batch_loci = std::atoi(need(i).c_str());
        else if (a == "--chunk-loci") o.chunk_loci = std::atoi(need(i).c_str());
        else if (a == "--history") o.history = std::atoi(need(i).c_str());
        else if (a == "--subspace-rank") o.subspace_rank = std::atoi(need(i).c_str());
        else if (a == "--max-iter") o.max_iter = std::atoi(need(i).c_str());
        else if (a == "--q-env") o.q_env = std::atoi(need(i).c_str());
        else if (a == "--formulation") {
            const std::string v = need(i);
            if (v == "auto") o.formulation = SDCA_FORM_AUTO;
            else if (v == "gram") o.formulation = SDCA_FORM_GRAM;
            else if (v == "primal") o.formulation = SDCA_FORM_PRIMAL;
            else die("unknown formulation \"" + v + "\"");
        } else if (a == "--precision") {
            const std::string v = need(i);
            if (v == "fp32") o.precision = SDCA_PREC_FP32;
            else if (v == "split-bf16") o.precision = SDCA_PREC_SPLIT_BF16;
            else die("unknown precision \"" + v + "\"");
        } else if (!a.empty() && a[0] == '-') {
            die("unrecognised option '" + a + "'\n" + kUsage);
        } else {
            o.alignments.push_back(a);
        }
    }
    if (o.alignments.empty()) {
        std::cerr << "SuperDCA ERROR: No alignment file specified!\n" << kUsage;
        std::exit(EXIT_FAILURE);
    }
    if (o.alignments.size() > 2)  // SuperDCA.cpp:155-159
        die("cannot handle more than 2 alignments (got " + std::to_string(o.alignments.size()) + " filenames)");
    if (o.norm_of_mean && (o.output_top > 0 || o.output_min_score > -HUGE_VAL))
        die("--norm-of-mean-scoring writes its own six-column output; --output-top and --output-min-score rank "
            "the default mean-of-norms scores only");
    if (o.chunk_loci < 0) die("--chunk-loci must be >= 0");
    return o;
}

struct Fasta {
    std::vector<std::string> names;
    std::vector<std::string> seqs;
};

Fasta read_fasta(const std::string &path)
{
    std::ifstream in(path);
    if (!in) die("Could not get alignment from input file \"" + path + "\"");
    Fasta f;
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        if (line[0] == '>') {
            f.names.push_back(line.substr(1));
            f.seqs.emplace_back();
        } else {
            if (f.seqs.empty()) die("\"" + path + "\" is not in FASTA format");
            f.seqs.back() += line;
        }
    }
    if (f.seqs.empty()) die("no sequences in \"" + path + "\"");
    for (const auto &s : f.seqs)
        if (s.size() != f.seqs[0].size()) die("sequences in \"" + path + "\" have different lengths");
    return f;
}

std::vector<long> read_index_list(const std::string &path, int base)
{
    std::ifstream in(path);
    if (!in) die("failed to parse list from file \"" + path + "\"");
    std::vector<long> v;
    std::string tok;
    while (in >> tok) {
        for (char &ch : tok)
            if (ch == ',') ch = ' ';
        std::istringstream ss(tok);
        long x;
        while (ss >> x) v.push_back(x - base);
    }
    return v;
}

void write_fasta(const std::string &path, const Fasta &f)
{
    std::ofstream out(path);
    for (size_t b = 0; b < f.seqs.size(); ++b) out << '>' << f.names[b] << '\n' << f.seqs[b] << '\n';
}

}  // namespace

sdca_opts solver_opts(const Options &o)
{
    sdca_opts so;
    sdca_default_opts(&so);
    so.device = o.device;
    so.precision = o.precision;
    so.formulation = o.formulation;
    so.batch_loci = o.batch_loci;  // 0: the library's default for the formulation it picks
    so.history = o.history;
    so.subspace_rank = o.subspace_rank;
    so.max_iter = o.max_iter;
    so.grad_threshold = o.gradient_threshold;
    so.j_threshold = o.coupling_threshold;
    so.verbose = o.verbose;
    return so;
}

// Two-alignment scan (SuperDCA.cpp:155-159, plmDCA.hpp:482-490): every locus of the first
// alignment against the loci of the second, weights from the second (plmDCA.hpp:538),
// "score r n" lines, not symmetrised, in <first>_scan.SuperDCA_couplings.<base>-based.all
// (plmDCA.hpp:598,672-681).
int run_scan(const Options &o)
{
    std::ostream &log = std::cout;
    if (!o.filterlist.empty() || !o.samplelist.empty() || o.out_filtered || o.out_filterlist || o.out_samplelist ||
        o.output_top > 0 || o.output_min_score > -HUGE_VAL)
        die("--filterlistfile, --samplelistfile, the --output-*-alignment options and the top-score output are "
            "not supported with two alignments");
    Fasta fa[2] = {read_fasta(o.alignments[0]), read_fasta(o.alignments[1])};
    const int B = (int)fa[0].seqs.size();
    if ((int)fa[1].seqs.size() != B) die("the two alignments hold different numbers of sequences");
    std::vector<uint8_t> st[2];
    std::vector<int32_t> tr[2];
    int32_t Ls[2] = {0, 0};
    for (int a = 0; a < 2; ++a) {
        const int L_in = (int)fa[a].seqs[0].size();
        std::vector<char> chars((size_t)B * L_in);
        for (int b = 0; b < B; ++b) std::memcpy(chars.data() + (size_t)b * L_in, fa[a].seqs[b].data(), L_in);
        st[a].resize((size_t)B * L_in);
        tr[a].resize(L_in);
        if (sdca_prepare_alignment(chars.data(), B, L_in, o.maf_threshold, o.gap_threshold, o.no_filter, o.q_env,
                                   st[a].data(), tr[a].data(), &Ls[a]) != SDCA_OK)
            die("alignment filtering failed");
        st[a].resize((size_t)B * Ls[a]);
        tr[a].resize(Ls[a]);
        if (o.verbose)
            log << "SuperDCA: alignment " << id_string(o.alignments[a]) << ": " << B << " sequences, " << Ls[a]
                << " of " << L_in << " loci pass the filter\n";
    }
    const int L1 = Ls[0], L2 = Ls[1];
    if (L1 < 1 || L2 < 1) {  // any non-empty second alignment (plmDCA.hpp:491)
        if (o.verbose) log << "plmDCA: there are not enough loci for plmDCA analysis\n";
        return EXIT_FAILURE;
    }
    std::vector<int32_t> loci;
    if (!o.locilist.empty()) {
        for (long v : read_index_list(o.locilist, o.input_base)) {
            if (v < 0 || v >= L1) die("loci list index out of range");
            loci.push_back((int32_t)v);
        }
    } else {
        const long end = o.end_locus < 0 ? L1 : std::min<long>(L1, o.end_locus + 1);
        for (long k = o.begin_locus; k < end; ++k) loci.push_back((int32_t)k);
    }
    if (loci.empty()) {
        if (o.verbose) log << "plmDCA: there are not enough loci for plmDCA analysis\n";
        return EXIT_FAILURE;
    }
    const sdca_opts so = solver_opts(o);
    std::vector<double> w(B, 1.0);
    if (!o.no_reweighting) {  // the second alignment's weights
        sdca_ctx *c2 = nullptr;
        if (sdca_ctx_create(&c2, st[1].data(), B, L2, o.q_env, o.q_env - 1, &so) != SDCA_OK)
            die(std::string("could not set up the GPU solve: ") + sdca_last_error());
        const int rc = sdca_reweight(c2, o.reweighting_threshold, nullptr, w.data(), nullptr);
        sdca_ctx_destroy(c2);
        if (rc != SDCA_OK) die(std::string("reweighting failed: ") + sdca_last_error());
    }
    const std::string id = id_string(o.alignments[0]);
    if (o.output_weights) {
        FILE *wf = fopen(unique_name(id + ".weights").c_str(), "w");  // plmDCA.hpp:550, also when scanning
        if (!wf) die("cannot write weights");
        for (double v : w) fprintf(wf, "%.8e\n", v);
        fclose(wf);
    }
    if (o.no_dca) return EXIT_SUCCESS;
    // the first alignment's loci, then the second's: loci [0, L1) are fitted on [L1, L1 + L2)
    const int L = L1 + L2;
    std::vector<uint8_t> both((size_t)B * L);
    for (int b = 0; b < B; ++b) {
        std::memcpy(both.data() + (size_t)b * L, st[0].data() + (size_t)b * L1, L1);
        std::memcpy(both.data() + (size_t)b * L + L1, st[1].data() + (size_t)b * L2, L2);
    }
    sdca_ctx *ctx = nullptr;
    if (sdca_ctx_create(&ctx, both.data(), B, L, o.q_env, o.q_env - 1, &so) != SDCA_OK)
        die(std::string("could not set up the GPU solve: ") + sdca_last_error());
    double lam[3];
    if (sdca_set_weights(ctx, w.data(), o.lambda_h, o.lambda_J, lam) != SDCA_OK ||
        sdca_set_scan_split(ctx, L1) != SDCA_OK)
        die(std::string("scan set-up failed: ") + sdca_last_error());
    if (o.verbose)
        log << "plmDCA: scan of " << loci.size() << " loci against " << L2 << " loci, n(effective)=" << lam[0]
            << "\n";
    const int K = (int)loci.size();
    std::vector<int32_t> t1(K);
    for (int k = 0; k < K; ++k) t1[k] = tr[0][loci[k]];
    const std::string out =
        unique_name(id + "_scan.SuperDCA_couplings." + std::to_string(o.output_base) + "-based.all");
    if (o.norm_of_mean) {
        // the q x q blocks of every (r, n), n over the second alignment, through a file-backed
        // store (plmDCA.hpp:340-345,625-638)
        const int q = o.q_env;
        const size_t n = (size_t)L * q * q + q;
        std::vector<int32_t> cols(L2);
        for (int u = 0; u < L2; ++u) cols[u] = L1 + u;
        sdca_store *store = nullptr;
        if (!o.no_coupling_output &&
            sdca_store_create(&store, unique_name(out + ".store").c_str(), K, L2, q * q) != SDCA_OK)
            die(std::string("cannot create the block store: ") + sdca_last_error());
        const int chunk = o.chunk_loci > 0 ? o.chunk_loci
                                           : (int)std::max<size_t>(1, std::min<size_t>(1024, (size_t(256) << 20) / (n * 8)));
        std::vector<double> xs((size_t)std::min(chunk, K) * n);
        std::vector<sdca_locus_stats> stats(K);
        for (int k0 = 0; k0 < K; k0 += chunk) {
            const int c = std::min(chunk, K - k0);
            const int rc = sdca_solve_loci(ctx, loci.data() + k0, c, nullptr, nullptr, xs.data(), stats.data() + k0);
            if (rc != SDCA_OK && rc != SDCA_E_NOCONV) die(std::string("solve failed: ") + sdca_last_error());
            if (rc == SDCA_E_NOCONV) std::cerr << "plmDCA warning: " << sdca_last_error() << "\n";
            for (int k = 0; k < c && store; ++k)
                if (sdca_store_put_blocks(store, k0 + k, xs.data() + (size_t)k * n, L, q, loci[k0 + k], cols.data()) !=
                    SDCA_OK)
                    die(std::string("cannot store coupling blocks: ") + sdca_last_error());
        }
        sdca_ctx_destroy(ctx);
        if (store) {
            if (o.verbose) log << "\nplmDCA: writing coupling values to file \"" << out << "\"\n";
            if (sdca_store_write_norm_of_mean_scan(store, out.c_str(), t1.data(), tr[1].data(), o.output_base, q,
                                                   q - 1) != SDCA_OK)
                die(std::string("could not write the couplings: ") + sdca_last_error());
            sdca_store_destroy(store);
        }
        return EXIT_SUCCESS;
    }
    std::vector<float> rows((size_t)K * L);
    std::vector<sdca_locus_stats> stats(K);
    const int rc = sdca_solve_loci(ctx, loci.data(), K, nullptr, rows.data(), nullptr, stats.data());
    if (rc != SDCA_OK && rc != SDCA_E_NOCONV) die(std::string("solve failed: ") + sdca_last_error());
    if (rc == SDCA_E_NOCONV) std::cerr << "plmDCA warning: " << sdca_last_error() << "\n";
    sdca_ctx_destroy(ctx);
    if (!o.no_coupling_output) {
        std::vector<float> S((size_t)K * L2);
        for (int k = 0; k < K; ++k)
            std::memcpy(S.data() + (size_t)k * L2, rows.data() + (size_t)k * L + L1, (size_t)L2 * 4);
        if (o.verbose) log << "\nplmDCA: writing coupling values to file \"" << out << "\"\n";
        if (sdca_write_scan_couplings(out.c_str(), S.data(), K, L2, t1.data(), tr[1].data(), o.output_base) !=
            SDCA_OK)
            die("could not write \"" + out + "\"");
    }
    return EXIT_SUCCESS;
}

int main(int argc, char **argv)
{
    const auto t_start = std::chrono::steady_clock::now();
    Options o = parse_args(argc, argv);
    if (o.alignments.size() == 2) return run_scan(o);
    std::ostream &log = std::cout;
    const std::string &path = o.alignments[0];
    if (o.verbose) log << "SuperDCA: get alignment from file \"" << path << "\"\n";
    Fasta fa = read_fasta(path);
    const std::string id = id_string(path);

    if (!o.samplelist.empty()) {  // keep the listed sample names, in file order of the alignment
        std::ifstream in(o.samplelist);
        if (!in) die("failed to parse sample list from \"" + o.samplelist + "\"");
        std::vector<std::string> keep;
        std::string nm;
        while (std::getline(in, nm)) {
            if (!nm.empty() && nm.back() == '\r') nm.pop_back();
            if (!nm.empty()) keep.push_back(nm[0] == '>' ? nm.substr(1) : nm);
        }
        Fasta g;
        for (size_t b = 0; b < fa.names.size(); ++b)
            if (std::find(keep.begin(), keep.end(), fa.names[b]) != keep.end()) {
                g.names.push_back(fa.names[b]);
                g.seqs.push_back(fa.seqs[b]);
            }
        if (g.seqs.empty()) die("sample list selects no sequences");
        fa = std::move(g);
        if (o.out_samplelist) write_fasta(unique_name(id + ".fasta"), fa);
    }
    std::vector<int32_t> column_of;  // input column of each column of fa (after the filter list)
    const int L_raw = (int)fa.seqs[0].size();
    for (int j = 0; j < L_raw; ++j) column_of.push_back(j);
    if (!o.filterlist.empty()) {
        auto cols = read_index_list(o.filterlist, o.input_base);
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        for (long cidx : cols)
            if (cidx < 0 || cidx >= L_raw) die("filter list index out of range");
        for (auto &s : fa.seqs) {
            std::string t;
            t.reserve(cols.size());
            for (long cidx : cols) t.push_back(s[cidx]);
            s.swap(t);
        }
        column_of.assign(cols.begin(), cols.end());
        if (o.out_filterlist) write_fasta(unique_name(id + ".fasta"), fa);
    }
    const int B = (int)fa.seqs.size();
    const int L_in = (int)fa.seqs[0].size();
    if (o.verbose) log << "SuperDCA: input alignment " << id << " has " << B << " sequences and " << L_in << " loci\n";
    std::vector<char> chars((size_t)B * L_in);
    for (int b = 0; b < B; ++b) std::memcpy(chars.data() + (size_t)b * L_in, fa.seqs[b].data(), L_in);
    std::vector<uint8_t> states((size_t)B * L_in);
    std::vector<int32_t> kept(L_in);
    int32_t L = 0;
    if (sdca_prepare_alignment(chars.data(), B, L_in, o.maf_threshold, o.gap_threshold, o.no_filter, o.q_env,
                               states.data(), kept.data(), &L) != SDCA_OK)
        die("alignment filtering failed");
    states.resize((size_t)B * L);
    std::vector<int32_t> translation(L);
    for (int k = 0; k < L; ++k) translation[k] = column_of[kept[k]];
    if (o.verbose)
        log << "SuperDCA: " << L << " loci pass the filter (maf >= " << o.maf_threshold << ", gap <= " << o.gap_threshold
            << ") and fit the " << o.q_env << "-state envelope\n";
    if (o.out_filtered) {
        Fasta g;
        g.names = fa.names;
        for (int b = 0; b < B; ++b) {
            std::string t;
            t.reserve(L);
            for (int k = 0; k < L; ++k) t.push_back(fa.seqs[b][kept[k]]);
            g.seqs.push_back(std::move(t));
        }
        write_fasta(unique_name(id + ".fasta"), g);
        std::ofstream lo(unique_name(id + ".loci"));
        for (int k = 0; k < L; ++k) lo << (k ? "," : "") << translation[k] + o.output_base;
        lo << "\n";
    }
    if (L < 2) {
        if (o.verbose) log << "plmDCA: there are not enough loci for plmDCA analysis (found " << L << " loci that qualify, but need at least 2)\n";
        return EXIT_FAILURE;
    }
    std::vector<int32_t> loci;
    if (!o.locilist.empty()) {
        for (long v : read_index_list(o.locilist, o.input_base)) {
            if (v < 0 || v >= L) die("loci list index out of range");
            loci.push_back((int32_t)v);
        }
    } else {
        const long end = o.end_locus < 0 ? L : std::min<long>(L, o.end_locus + 1);
        for (long k = o.begin_locus; k < end; ++k) loci.push_back((int32_t)k);
    }
    if (loci.size() < 2) {
        if (o.verbose) log << "plmDCA: there are not enough loci for plmDCA analysis (need at least 2)\n";
        return EXIT_FAILURE;
    }

    sdca_opts so = solver_opts(o);
    sdca_ctx *ctx = nullptr;
    if (sdca_ctx_create(&ctx, states.data(), B, L, o.q_env, o.q_env - 1, &so) != SDCA_OK)
        die(std::string("could not set up the GPU solve: ") + sdca_last_error());
    std::vector<double> w(B, 1.0);
    if (!o.no_reweighting) {
        if (o.verbose) log << "\nplmDCA: calculate sequence weights\n";
        if (sdca_reweight(ctx, o.reweighting_threshold, nullptr, w.data(), nullptr) != SDCA_OK)
            die(std::string("reweighting failed: ") + sdca_last_error());
    }
    double lam[3];
    if (sdca_set_weights(ctx, w.data(), o.lambda_h, o.lambda_J, lam) != SDCA_OK)
        die(std::string("weights rejected: ") + sdca_last_error());
    if (o.verbose)
        log << "plmDCA: L=" << L << " n=" << B << " n(effective)=" << lam[0] << "\n"
            << "plmDCA: lambda_J=" << lam[2] << " lambda_h=" << lam[1] << " (absolute)\n";
    if (o.output_weights) {
        FILE *wf = fopen(unique_name(id + ".weights").c_str(), "w");
        if (!wf) die("cannot write weights");
        for (double v : w) fprintf(wf, "%.8e\n", v);
        fclose(wf);
    }
    if (o.no_dca) {
        if (o.verbose) log << "plmDCA: exit on users request without performing DCA\n\n";
        sdca_ctx_destroy(ctx);
        return EXIT_SUCCESS;
    }
    if (o.verbose) log << "\nplmDCA: let's learn!\n";
    const int K = (int)loci.size();
    const int q = o.q_env;
    const size_t n = (size_t)L * q * q + q;  // parameters per locus (reference layout)
    const bool want_top = o.output_top > 0 || o.output_min_score > -HUGE_VAL;
    const bool want_all = !o.no_coupling_output;
    const bool nom = o.norm_of_mean;
    const bool keep = want_all || want_top;
    const std::string stem = id + ".SuperDCA_couplings." + std::to_string(o.output_base) + "-based";
    // the scores (or the coupling blocks) of the listed loci, row and column = list position
    // (plmDCA.hpp:684-691), go to a file-backed store chunk by chunk
    sdca_store *store = nullptr;
    if (keep) {
        const std::string sp = unique_name(stem + ".store");
        if (sdca_store_create(&store, sp.c_str(), K, K, nom ? q * q : 1) != SDCA_OK)
            die(std::string("cannot create the score store: ") + sdca_last_error());
    }
    const int chunk = o.chunk_loci > 0 ? o.chunk_loci
                                       : (nom ? (int)std::max<size_t>(1, std::min<size_t>(1024, (size_t(256) << 20) / (n * 8)))
                                              : 4096);
    std::vector<float> rows(keep && !nom ? (size_t)std::min(chunk, K) * L : 0);
    std::vector<float> band(keep && !nom ? (size_t)std::min(chunk, K) * K : 0);
    std::vector<double> xs(keep && nom ? (size_t)std::min(chunk, K) * n : 0);
    std::vector<sdca_locus_stats> st(K);
    const auto t_solve = std::chrono::steady_clock::now();
    int nwarn = 0;
    for (int k0 = 0; k0 < K; k0 += chunk) {
        const int c = std::min(chunk, K - k0);
        const int rc = sdca_solve_loci(ctx, loci.data() + k0, c, nullptr, rows.empty() ? nullptr : rows.data(),
                                       xs.empty() ? nullptr : xs.data(), st.data() + k0);
        if (rc != SDCA_OK && rc != SDCA_E_NOCONV) die(std::string("solve failed: ") + sdca_last_error());
        if (rc == SDCA_E_NOCONV) ++nwarn;
        if (!rows.empty()) {
            for (int k = 0; k < c; ++k)
                for (int u = 0; u < K; ++u) band[(size_t)k * K + u] = rows[(size_t)k * L + loci[u]];
            if (sdca_store_put_rows(store, k0, c, band.data()) != SDCA_OK)
                die(std::string("cannot store scores: ") + sdca_last_error());
        }
        if (!xs.empty())
            for (int k = 0; k < c; ++k)
                if (sdca_store_put_blocks(store, k0 + k, xs.data() + (size_t)k * n, L, q, loci[k0 + k], loci.data()) !=
                    SDCA_OK)
                    die(std::string("cannot store coupling blocks: ") + sdca_last_error());
        if (o.verbose && K > chunk) log << "plmDCA: " << k0 + c << " / " << K << " loci solved\n";
    }
    const double solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_solve).count();
    if (nwarn) {
        int nc = 0;
        for (const auto &x : st) nc += x.status == SDCA_ST_ITER_CAP || x.status == SDCA_ST_LS_FAIL;
        std::cerr << "plmDCA warning: " << nc << " loci stopped without reaching the gradient threshold\n";
    }
    if (o.verbose) {
        for (int k = 0; k < K; ++k)
            log << "  " << k + 1 << " / " << K << " (out of " << L << ") locus=" << translation[loci[k]] + 1
                << " fval=" << std::scientific << st[k].fval << " gnorm=" << st[k].gnorm << std::defaultfloat
                << " nfeval=" << st[k].nfeval << " iter=" << st[k].iter << "\n";
        log << "plmDCA: solved " << K << " loci in " << solve_s << " s (" << K / solve_s << " loci/s)\n";
    }
    if (store) {
        std::vector<int32_t> tr(K);
        for (int k = 0; k < K; ++k) tr[k] = translation[loci[k]];
        const std::string out_all = want_all ? unique_name(stem + ".all") : std::string();
        const std::string out_top = want_top ? unique_name(stem + ".top") : std::string();
        if (o.verbose && want_all) log << "\nplmDCA: writing coupling values to file \"" << out_all << "\"\n";
        if (o.verbose && want_top) log << "plmDCA: writing the largest coupling values to file \"" << out_top << "\"\n";
        const int rc = nom ? sdca_store_write_norm_of_mean(store, out_all.c_str(), tr.data(), o.output_base, q, q - 1, 0)
                           : sdca_store_write_couplings(store, want_all ? out_all.c_str() : nullptr,
                                                        want_top ? out_top.c_str() : nullptr, tr.data(), o.output_base,
                                                        o.output_top, o.output_min_score, 0);
        if (rc != SDCA_OK) die(std::string("could not write the couplings: ") + sdca_last_error());
        sdca_store_destroy(store);
    }
    sdca_ctx_destroy(ctx);
    if (o.verbose) {
        log << "\nplmDCA: analysis completed\n";
        log << "SuperDCA: total time "
            << std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count() << " s\n";
    }
    return EXIT_SUCCESS;
}
