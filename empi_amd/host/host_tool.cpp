// empi-host-tool: host-side surfaces without a device, for tests and tools.
//   conf <config>           parsed entries, "key<TAB>value" per line (map order)
//   subset <text> <max>     parsed integer subset
//   book <desc> <out.b>     writes a book from a text description:
//                           "fs C N epochs", then per epoch C lines of N
//                           samples, then per channel "count" and count
//                           lines of 6 parameters
//   readbook <book.b>       prints a book as text
//   read <signal> <channels> <selected> whole|all|selected [epochSize] [epochs]
//                           reads the file with the signal readers until an
//                           empty read: "epoch k samples n", then one line of
//                           samples per selected channel; lists are 1-based,
//                           comma-separated
// Errors print "ERROR<TAB><code>" and exit 1.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>

#include "book.hpp"
#include "config.hpp"
#include "signal_reader.hpp"

using namespace empi::host;

static int book_from_text(const char* desc, const char* out) {
    std::ifstream in(desc);
    double fs;
    int C, N, E;
    if (!(in >> fs >> C >> N >> E)) throw Error("badDescription");
    BookWriter w(out);
    for (int e = 1; e <= E; ++e) {
        Signal s;
        s.fs = fs;
        s.channels = C;
        s.samples = N;
        s.data.resize(static_cast<size_t>(C) * N);
        for (double& v : s.data) in >> v;
        ChannelAtoms atoms(static_cast<size_t>(C));
        for (int c = 0; c < C; ++c) {
            int count;
            in >> count;
            for (int i = 0; i < count; ++i) {
                Atom a;
                for (double& p : a.params) in >> p;
                atoms[c].push_back(a);
            }
        }
        if (!in) throw Error("badDescription");
        w.write(e, s, atoms);
    }
    w.close();
    return 0;
}

static std::vector<int> int_list(const std::string& text) {
    std::vector<int> out;
    size_t pos = 0;
    while (pos <= text.size()) {
        const size_t end = text.find(',', pos);
        const std::string tok = text.substr(pos, end == std::string::npos ? std::string::npos : end - pos);
        if (!tok.empty()) out.push_back(std::atoi(tok.c_str()));
        if (end == std::string::npos) break;
        pos = end + 1;
    }
    return out;
}

static int read_signal(int argc, char** argv) {
    if (argc < 6) return 2;
    const std::string mode = argv[5];
    std::unique_ptr<SignalReader> r;
    if (mode == "whole") {
        r.reset(new SignalReaderForWholeSignal(argv[2]));
    } else if (mode == "all" && argc >= 7) {
        r.reset(new SignalReaderForAllEpochs(argv[2], std::atoi(argv[6])));
    } else if (mode == "selected" && argc >= 8) {
        r.reset(new SignalReaderForSelectedEpochs(argv[2], std::atoi(argv[6]), int_list(argv[7])));
    } else {
        return 2;
    }
    r->channelCount = std::atoi(argv[3]);
    r->selectedChannels = int_list(argv[4]);
    for (int k = 1;; ++k) {
        Signal s = r->read();
        if (s.samples == 0) break;
        std::printf("epoch %d samples %d\n", k, s.samples);
        for (int c = 0; c < s.channels; ++c) {
            for (int n = 0; n < s.samples; ++n)
                std::printf(n ? " %.9g" : "%.9g", s.data[static_cast<size_t>(c) * s.samples + n]);
            std::printf("\n");
        }
        std::fflush(stdout);
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s conf|subset|book|readbook ...\n", argv[0]);
        return 2;
    }
    const std::string cmd = argv[1];
    try {
        if (cmd == "conf") {
            Configuration c;
            c.parse(argv[2]);
            for (const auto& kv : c.entries()) std::cout << kv.first << '\t' << kv.second << '\n';
        } else if (cmd == "subset" && argc >= 4) {
            for (int v : parse_integer_subset(argv[2], std::atoi(argv[3]))) std::cout << v << ' ';
            std::cout << '\n';
        } else if (cmd == "book" && argc >= 4) {
            return book_from_text(argv[2], argv[3]);
        } else if (cmd == "read") {
            return read_signal(argc, argv);
        } else if (cmd == "readbook") {
            Book b = read_book(argv[2]);
            std::printf("fs %.9g channels %d epochs %zu\n", b.fs, b.channels, b.epochs.size());
            for (const BookEpoch& e : b.epochs) {
                std::printf("epoch %d samples %d\n", e.epoch, e.samples);
                for (size_t c = 0; c < e.atoms.size(); ++c)
                    for (const auto& a : e.atoms[c]) {
                        std::printf("%zu", c + 1);
                        for (float p : a) std::printf(" %.9g", p);
                        std::printf("\n");
                    }
            }
        } else {
            return 2;
        }
    } catch (const Error& e) {
        std::cout << "ERROR\t" << e.what() << std::endl;
        return 1;
    }
    return 0;
}
