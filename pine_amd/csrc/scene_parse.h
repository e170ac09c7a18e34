// pine's scene-file grammar and parameter tree, restated in C++ for the
// drop-in host side (src/util/parser.cpp:45-110, src/util/parameters.cpp).
// Numbers go through pine's own stof/stoi (src/pstd/string.cpp:36-121): they
// accumulate digits without correct rounding and accept no exponent, and scene
// values must round exactly as in pine to reproduce its geometry.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace pg {

float pine_stof(const std::string& s);
int pine_stoi(const std::string& s);
void pine_stofs(const std::string& s, float* out, int n);
void pine_stois(const std::string& s, int* out, int n);

// Parameters (src/util/parameters.h:14-86): a value map plus named lists of
// sub-blocks.  A typed block `key: Type{...}` stores its type under "@".
struct Params {
    std::map<std::string, std::string> values;
    std::map<std::string, std::vector<Params>> subsets;

    bool has_value(const std::string& k) const { return values.count(k) != 0; }
    bool has_subset(const std::string& k) const { return subsets.count(k) || has_value(k); }
    // operator[] semantics of parameters.cpp:26-41 (a plain value becomes a
    // sub-block whose "@" is that value); returns an empty block if absent.
    const Params& sub(const std::string& k) const;
    const std::vector<Params>& all(const std::string& k) const;

    std::string get_string(const std::string& k, const std::string& def = "") const;
    bool try_string(const std::string& k, std::string& out) const;
    float get_float(const std::string& k, float def = 0.0f) const;
    int get_int(const std::string& k, int def = 0) const;
    bool get_bool(const std::string& k, bool def = false) const;
    bool try_vec3(const std::string& k, float out[3]) const;
    void get_vec3(const std::string& k, float out[3], const float* def = nullptr) const;
    void get_vec2i(const std::string& k, int out[2], const int* def = nullptr) const;
    void get_vec2(const std::string& k, float out[2], const float* def = nullptr) const;

  private:
    mutable std::map<std::string, std::vector<Params>> lazy_;
};

// Parses pine scene text; throws std::runtime_error on malformed input (pine
// aborts through CHECK in the same places, parser.cpp:55,61,79).
Params parse_scene_text(const std::string& text);

}  // namespace pg
