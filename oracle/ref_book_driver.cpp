// Text driver around the reference's BookWriter (src/io.cpp:87-168), used
// only to pin the host book writer; same description format as
// empi-host-tool's "book" command: argv[1] description, argv[2] output.
#include <fstream>
#include <iostream>
#include "base.hpp"
#include "io.hpp"

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    std::ifstream in(argv[1]);
    double fs;
    int C, N, E;
    if (!(in >> fs >> C >> N >> E)) return 2;
    BookWriter w(argv[2]);
    for (int e = 1; e <= E; ++e) {
        MultiSignal s;
        s.channels.resize(C);
        for (int c = 0; c < C; ++c) {
            s.channels[c].freqSampling = fs;
            s.channels[c].samples.resize(N);
            for (int n = 0; n < N; ++n) in >> s.channels[c].samples[n];
        }
        MultiChannelResult r(C);
        for (int c = 0; c < C; ++c) {
            int count;
            in >> count;
            for (int i = 0; i < count; ++i) {
                Atom a;
                a.type = ATOM_GABOR;
                a.params.resize(6);
                for (double& p : a.params) in >> p;
                r[c].push_back(a);
            }
        }
        w.write(e, s, r);
    }
    w.close();
    return 0;
}
