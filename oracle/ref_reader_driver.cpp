// Text driver around the reference's signal readers (src/io.cpp:9-85), used
// only to pin the host readers; same arguments and output as empi-host-tool's
// "read" command:
//   argv: <signal> <channels> <selected> whole|all|selected [epochSize] [epochs]
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>
#include <vector>
#include "base.hpp"
#include "io.hpp"

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

int main(int argc, char** argv) {
    if (argc < 5) return 2;
    const std::string mode = argv[4];
    try {
        std::unique_ptr<SignalReader> r;
        if (mode == "whole") r.reset(new SignalReaderForWholeSignal(argv[1]));
        else if (mode == "all" && argc >= 6) r.reset(new SignalReaderForAllEpochs(argv[1], std::atoi(argv[5])));
        else if (mode == "selected" && argc >= 7)
            r.reset(new SignalReaderForSelectedEpochs(argv[1], std::atoi(argv[5]), int_list(argv[6])));
        else return 2;
        r->channelCount = std::atoi(argv[2]);
        r->selectedChannels = int_list(argv[3]);
        for (int k = 1;; ++k) {
            MultiSignal s = r->read();
            const int n = static_cast<int>(s.channels[0].samples.size());
            if (!n) break;
            std::printf("epoch %d samples %d\n", k, n);
            for (const auto& ch : s.channels) {
                for (int i = 0; i < n; ++i) std::printf(i ? " %.9g" : "%.9g", static_cast<double>(ch.samples[i]));
                std::printf("\n");
            }
            std::fflush(stdout);
        }
    } catch (const std::exception& e) {
        std::cout << "ERROR\t" << e.what() << std::endl;
        return 1;
    }
    return 0;
}
