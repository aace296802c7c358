// Command-line parsing and output formats of the `visqol` tool (cli.cpp),
// header-only so a test program can check them without a device.  Flags
// and formats follow the reference CLI (main.zig:10-33 flags, 182-196 batch
// CSV, 245-257 result printing), which follows google/visqol's.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <charconv>
#include <string>
#include <utility>
#include <vector>

namespace vq_cli {

struct Args {
    std::string ref, deg, batch_csv, results_csv, model;
    bool speech = false, lattice = true, unscaled = false, no_global = false, no_realign = false;
    bool verbose = false, help = false;
    int32_t radius = 60;
    uint32_t backend = 0;     // VISQOL_BACKEND_CPU, the reference's default
    uint32_t precision = 3;   // VISQOL_PRECISION_AUTO
    uint32_t threads = 0;     // 0: all cores
};

inline const char* usage() {
    return "usage: visqol --reference_file REF.wav --degraded_file DEG.wav [options]\n"
           "       visqol --batch_input_csv PAIRS.csv [--results_csv OUT.csv] [options]\n"
           "\n"
           "  --similarity_to_quality_model F   SVR model (audio) or TFLite lattice model\n"
           "                                    (speech); default: the bundled one\n"
           "  --use_speech_mode                 speech mode (21 bands, VAD)\n"
           "  --use_lattice_model[=true|false]   speech mode: TF-Lattice mapping (default\n"
           "                                    true); false (or --nouse_lattice_model)\n"
           "                                    selects the exponential mapping\n"
           "  --use_unscaled_speech_mos_mapping exponential mapping without the scale that\n"
           "                                    maps NSIM 1 to MOS 5\n"
           "  --search_window_radius N          DP search radius in frames (default 60)\n"
           "  --disable_global_alignment\n"
           "  --disable_realignment\n"
           "  --verbose                         print per-band and per-patch detail\n"
           "  --backend cpu|hip                 cpu (default) and hip both score on the\n"
           "                                    gfx950 device; cpu always in exact f64\n"
           "  --threads N                       WAV decoding threads (default: all cores)\n"
           "  --precision exact|balanced|fast   exact: f64, the reference CPU path's scores\n"
           "                                    bit for bit; balanced and fast: the\n"
           "                                    throughput mode (|dMOS| <= 1e-6)\n"
           "                                    (default: exact)\n";
}

// Parse status: 0 parsed; 1 a malformed value (the parse error's name in
// err); 2 an unknown flag (err); 3 an unknown backend or precision (the
// message in err).  All but 2 exit 1.
inline int parse_args(int argc, const char* const* argv, Args& a, std::string& err) {
    for (int i = 1; i < argc; i++) {
        const std::string f = argv[i];
        auto next = [&](std::string& dst) -> bool {
            if (i + 1 >= argc) {
                err = "MissingArgument";
                return false;
            }
            dst = argv[++i];
            return true;
        };
        std::string v;
        if (f == "--reference_file") {
            if (!next(a.ref)) return 1;
        } else if (f == "--degraded_file") {
            if (!next(a.deg)) return 1;
        } else if (f == "--batch_input_csv") {
            if (!next(a.batch_csv)) return 1;
        } else if (f == "--results_csv") {
            if (!next(a.results_csv)) return 1;
        } else if (f == "--similarity_to_quality_model") {
            if (!next(a.model)) return 1;
        } else if (f == "--use_speech_mode") {
            a.speech = true;
        } else if (f == "--use_lattice_model" || f == "--use_lattice_model=true") {
            a.lattice = true;
        } else if (f == "--nouse_lattice_model" || f == "--use_lattice_model=false") {
            a.lattice = false;
        } else if (f == "--use_unscaled_speech_mos_mapping") {
            a.unscaled = true;
        } else if (f == "--search_window_radius") {
            if (!next(v)) return 1;
            int32_t r = 0;
            const auto res = std::from_chars(v.data(), v.data() + v.size(), r);
            if (res.ec != std::errc() || res.ptr != v.data() + v.size()) {
                err = res.ec == std::errc::result_out_of_range ? "Overflow" : "InvalidCharacter";
                return 1;
            }
            a.radius = r;
        } else if (f == "--disable_global_alignment") {
            a.no_global = true;
        } else if (f == "--disable_realignment") {
            a.no_realign = true;
        } else if (f == "--verbose") {
            a.verbose = true;
        } else if (f == "--backend") {
            if (!next(v)) return 1;
            if (v == "cpu") a.backend = 0;
            else if (v == "metal") a.backend = 1;
            else if (v == "cuda") a.backend = 2;
            else if (v == "vulkan") a.backend = 3;
            else if (v == "hip") a.backend = 4;
            else {
                err = "unknown backend";
                return 3;
            }
        } else if (f == "--threads") {
            if (!next(v)) return 1;
            uint32_t t = 0;
            const auto res = std::from_chars(v.data(), v.data() + v.size(), t);
            if (res.ec != std::errc() || res.ptr != v.data() + v.size()) {
                err = res.ec == std::errc::result_out_of_range ? "Overflow" : "InvalidCharacter";
                return 1;
            }
            a.threads = t;
        } else if (f == "--precision") {
            if (!next(v)) return 1;
            if (v == "exact") a.precision = 0;
            else if (v == "fast") a.precision = 1;
            else if (v == "balanced") a.precision = 2;
            else {
                err = "unknown precision";
                return 3;
            }
        } else if (f == "-h" || f == "--help") {
            a.help = true;
            return 0;
        } else {
            err = f;
            return 2;
        }
    }
    return 0;
}

inline std::string trim_spaces(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && s[b] == ' ') b++;
    while (e > b && s[e - 1] == ' ') e--;
    return s.substr(b, e - b);
}

// The batch CSV (main.zig:142-157): lines split on CR and LF (empty lines
// skipped), the first two comma-separated fields trimmed of spaces; a line
// with fewer than two fields is skipped; a first row "reference,..." is the
// header.
inline std::vector<std::pair<std::string, std::string>> parse_csv(const std::string& text) {
    std::vector<std::pair<std::string, std::string>> out;
    bool first = true;
    size_t pos = 0;
    while (pos < text.size()) {
        size_t end = text.find_first_of("\r\n", pos);
        if (end == std::string::npos) end = text.size();
        const std::string line = text.substr(pos, end - pos);
        pos = end + 1;
        if (line.empty()) continue;
        const size_t c1 = line.find(',');
        if (c1 == std::string::npos) continue;
        const size_t c2 = line.find(',', c1 + 1);
        const std::string r = trim_spaces(line.substr(0, c1));
        const std::string d = trim_spaces(line.substr(c1 + 1, c2 == std::string::npos ? std::string::npos : c2 - c1 - 1));
        if (first && r == "reference") {
            first = false;
            continue;
        }
        first = false;
        out.emplace_back(r, d);
    }
    return out;
}

// A double as the shortest decimal that reads back to it, without an
// exponent (the results CSV's score column).
inline std::string shortest_decimal(double x) {
    char buf[512];
    const auto res = std::to_chars(buf, buf + sizeof buf, x, std::chars_format::fixed);
    return std::string(buf, res.ptr);
}

}  // namespace vq_cli
