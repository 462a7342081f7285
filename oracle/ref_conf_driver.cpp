// Text driver around the reference's Configuration (src/conf.cpp), used
// only to pin the host config parser: prints parsed entries of argv[1] as
// "key<TAB>value" for the given keys (argv[2..]) that are present.
#include <iostream>
#include "base.hpp"
#include "conf.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    Configuration c;
    try {
        c.parse(argv[1]);
        for (int i = 2; i < argc; ++i)
            if (c.has(argv[i])) std::cout << argv[i] << '\t' << c.at(argv[i]) << '\n';
    } catch (const Exception& e) {
        std::cout << "ERROR\t" << e.what() << std::endl;
        return 1;
    }
    return 0;
}
