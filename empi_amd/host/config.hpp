// empi configuration file (src/conf.cpp:10-48, inc/conf.hpp:12-19).
//
// One "key<blanks>value" per line; lines whose first non-blank is '#' and
// lines without a value are skipped; the value runs to the end of the line
// (trailing blanks kept); a repeated key keeps the last value.
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace empi {
namespace host {

// Error with a reference camel-case code (inc/base.hpp:69-73).
class Error : public std::runtime_error {
public:
    explicit Error(const std::string& code) : std::runtime_error(code) {}
};

class Configuration {
public:
    void parse(const std::string& path);           // throws couldNotOpenConfigFile
    bool has(const std::string& key) const;
    const std::string& at(const std::string& key) const;   // throws missingConfigurationEntry:<key>
    const std::map<std::string, std::string>& entries() const { return values_; }

private:
    std::map<std::string, std::string> values_;
};

// "1-3,5,8-10" (src/empi.cpp:22-47): ranges and single numbers in [1, max].
std::vector<int> parse_integer_subset(const std::string& text, int max);

}  // namespace host
}  // namespace empi
