#include "config.hpp"

#include <climits>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace empi {
namespace host {

void Configuration::parse(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw Error("couldNotOpenConfigFile");
    values_.clear();
    const char* blanks = " \t";
    std::string line;
    while (std::getline(in, line)) {
        const size_t key_begin = line.find_first_not_of(blanks);
        if (key_begin == std::string::npos || line[key_begin] == '#') continue;
        const size_t key_end = line.find_first_of(blanks, key_begin + 1);
        if (key_end == std::string::npos) continue;
        const size_t value_begin = line.find_first_not_of(blanks, key_end + 1);
        if (value_begin == std::string::npos) continue;
        std::string key = line.substr(key_begin, key_end - key_begin);
        std::string value = line.substr(value_begin);
        if (!key.empty() && !value.empty()) values_[key] = value;
    }
}

bool Configuration::has(const std::string& key) const { return values_.count(key) != 0; }

const std::string& Configuration::at(const std::string& key) const {
    auto it = values_.find(key);
    if (it == values_.end()) throw Error("missingConfigurationEntry:" + key);
    return it->second;
}

std::vector<int> parse_integer_subset(const std::string& text, int max) {
    std::vector<int> out;
    std::stringstream ss(text);
    std::string part;
    while (std::getline(ss, part, ',')) {
        int a = 0, b = 0;
        const int n = std::sscanf(part.c_str(), "%d-%d", &a, &b);
        if (n == 2 && a >= 1 && b <= max) {
            for (int i = a; i <= b; ++i) out.push_back(i);
        } else if (n == 1 && a >= 1 && a <= max) {
            out.push_back(a);
        } else {
            throw Error("invalidSubsetSpecification");
        }
    }
    if (out.empty()) throw Error("emptySubsetSpecification");
    if (out.size() > static_cast<size_t>(INT_MAX)) throw Error("tooLargeSubsetSpecification");
    return out;
}

}  // namespace host
}  // namespace empi
