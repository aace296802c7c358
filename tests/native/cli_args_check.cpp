// Prints what the visqol tool's parsers make of their inputs, for
// tests/test_cli_cpu.py: `args ARG...` prints the parse status and fields;
// `csv FILE` the pairs of a batch CSV; `num X...` each double's results-CSV
// text.
#include <stdio.h>
#include <stdlib.h>

#include <fstream>
#include <sstream>
#include <string>

#include "cli_args.h"

int main(int argc, char** argv) {
    using namespace vq_cli;
    const std::string what = argc > 1 ? argv[1] : "";
    if (what == "args") {
        Args a;
        std::string err;
        const int st = parse_args(argc - 1, (const char* const*)argv + 1, a, err);
        printf("status %d\nerr %s\nref %s\ndeg %s\ncsv %s\nresults %s\nmodel %s\n", st, err.c_str(), a.ref.c_str(),
               a.deg.c_str(), a.batch_csv.c_str(), a.results_csv.c_str(), a.model.c_str());
        printf("speech %d\nlattice %d\nunscaled %d\nno_global %d\nno_realign %d\nverbose %d\nhelp %d\n", a.speech,
               a.lattice, a.unscaled, a.no_global, a.no_realign, a.verbose, a.help);
        printf("radius %d\nbackend %u\nprecision %u\nthreads %u\n", a.radius, a.backend, a.precision, a.threads);
        return 0;
    }
    if (what == "csv") {
        std::ifstream f(argv[2], std::ios::binary);
        std::ostringstream ss;
        ss << f.rdbuf();
        for (const auto& p : parse_csv(ss.str())) printf("[%s]|[%s]\n", p.first.c_str(), p.second.c_str());
        return 0;
    }
    if (what == "num") {
        for (int i = 2; i < argc; i++) printf("%s\n", shortest_decimal(strtod(argv[i], nullptr)).c_str());
        return 0;
    }
    return 2;
}
