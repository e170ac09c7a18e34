// pine's scene text parser and parameter getters, restated.
#include "scene_parse.h"

#include <stdexcept>

namespace pg {

// pstd/string.cpp:36-48
int pine_stoi(const std::string& str) {
    int number = 0;
    bool neg = false;
    for (size_t j = 0; j < str.size(); j++) {
        if (str[j] == '.')
            break;
        if (j == 0 && str[j] == '-')
            neg = true;
        else
            number = number * 10 + str[j] - '0';
    }
    return neg ? -number : number;
}

// pstd/string.cpp:50-68: digits accumulate as number*10 + d, decimals as
// d*scale with scale *= 0.1f; no exponent.
float pine_stof(const std::string& str) {
    float number = 0.0f;
    bool neg = false, decimal = false;
    float scale = 0.1f;
    for (size_t j = 0; j < str.size(); j++) {
        if (j == 0 && str[j] == '-')
            neg = true;
        else if (!decimal && str[j] == '.')
            decimal = true;
        else if (!decimal)
            number = number * 10 + str[j] - '0';
        else {
            number += (str[j] - '0') * scale;
            scale *= 0.1f;
        }
    }
    return neg ? -number : number;
}

static bool is_number(char c) { return c >= '0' && c <= '9'; }

// pstd/string.cpp:71-96
void pine_stois(const std::string& str, int* ptr, int N) {
    int dim = 0, start = -1;
    for (int i = 0; i < (int)str.size(); i++) {
        if (is_number(str[i]) || str[i] == '-') {
            if (start == -1)
                start = i;
        } else if (start != -1) {
            ptr[dim++] = pine_stoi(str.substr(start, i - start));
            start = -1;
            if (dim == N)
                return;
        }
    }
    if (start != -1)
        ptr[dim++] = pine_stoi(str.substr(start));
    if (dim == 0) {
        ptr[0] = 0;
        dim++;
    }
    for (int i = dim; i < N; i++)
        ptr[i] = ptr[dim - 1];
}

// pstd/string.cpp:98-121
void pine_stofs(const std::string& str, float* ptr, int N) {
    int dim = 0, start = -1;
    for (int i = 0; i < (int)str.size(); i++) {
        if (is_number(str[i]) || str[i] == '-' || str[i] == '.') {
            if (start == -1)
                start = i;
        } else if (start != -1) {
            ptr[dim++] = pine_stof(str.substr(start, i - start));
            start = -1;
            if (dim == N)
                return;
        }
    }
    if (start != -1)
        ptr[dim++] = pine_stof(str.substr(start));
    if (dim == 0) {
        ptr[0] = 0;
        dim++;
    }
    for (int i = dim; i < N; i++)
        ptr[i] = ptr[dim - 1];
}

// ---- Params (parameters.cpp) -------------------------------------------------

const Params& Params::sub(const std::string& k) const {
    auto it = subsets.find(k);
    if (it != subsets.end() && !it->second.empty())
        return it->second.back();
    auto& cache = lazy_[k];
    if (cache.empty()) {
        cache.resize(1);
        if (has_value(k))
            cache[0].values["@"] = values.at(k);
    }
    return cache[0];
}

const std::vector<Params>& Params::all(const std::string& k) const {
    static const std::vector<Params> none;
    auto it = subsets.find(k);
    return it == subsets.end() ? none : it->second;
}

bool Params::try_string(const std::string& k, std::string& out) const {
    auto it = values.find(k);
    if (it == values.end()) {
        if (k == "type" && has_value("@")) {
            out = values.at("@");
            return true;
        }
        return false;
    }
    out = it->second;
    return true;
}

std::string Params::get_string(const std::string& k, const std::string& def) const {
    std::string s;
    return try_string(k, s) ? s : def;
}

float Params::get_float(const std::string& k, float def) const {
    auto it = values.find(k);
    return it == values.end() ? def : pine_stof(it->second);
}

int Params::get_int(const std::string& k, int def) const {
    auto it = values.find(k);
    return it == values.end() ? def : pine_stoi(it->second);
}

bool Params::get_bool(const std::string& k, bool def) const {
    auto it = values.find(k);
    if (it == values.end())
        return def;
    std::string s = it->second;
    for (auto& c : s)
        if (c >= 'A' && c <= 'Z')
            c = char(c + ('a' - 'A'));
    if (s == "true")
        return true;
    if (s == "false")
        return false;
    return pine_stoi(s) != 0;
}

bool Params::try_vec3(const std::string& k, float out[3]) const {
    auto it = values.find(k);
    if (it == values.end())
        return false;
    pine_stofs(it->second, out, 3);
    return true;
}

void Params::get_vec3(const std::string& k, float out[3], const float* def) const {
    if (try_vec3(k, out))
        return;
    if (!def)
        throw std::runtime_error("missing parameter \"" + k + "\"");
    for (int i = 0; i < 3; i++)
        out[i] = def[i];
}

void Params::get_vec2(const std::string& k, float out[2], const float* def) const {
    auto it = values.find(k);
    if (it != values.end()) {
        pine_stofs(it->second, out, 2);
        return;
    }
    if (!def)
        throw std::runtime_error("missing parameter \"" + k + "\"");
    out[0] = def[0];
    out[1] = def[1];
}

void Params::get_vec2i(const std::string& k, int out[2], const int* def) const {
    auto it = values.find(k);
    if (it != values.end()) {
        pine_stois(it->second, out, 2);
        return;
    }
    if (!def)
        throw std::runtime_error("missing parameter \"" + k + "\"");
    out[0] = def[0];
    out[1] = def[1];
}

// ---- parser (parser.cpp:9-101) ------------------------------------------------
namespace {

bool is_letter(char s) { return (s >= 'A' && s <= 'z') || (s >= '0' && s <= '9') || s == '_'; }
bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\f'; }

struct View {
    const std::string* s;
    size_t pos;
    size_t size() const { return s->size() - pos; }
    char at(size_t i) const { return i < size() ? (*s)[pos + i] : '\0'; }
    std::string str(size_t b, size_t n) const { return s->substr(pos + b, n); }
    void skip(size_t n) { pos = n > size() ? s->size() : pos + n; }
};

template <typename F>
bool first_of(const View& v, F&& f, size_t& out) {
    for (size_t i = 0; i < v.size(); i++)
        if (f(v.at(i))) {
            out = i;
            return true;
        }
    return false;
}

// EscapeSpace: a "//" comment is only recognised right after whitespace.
void escape_space(View& v) {
    while (v.size() && is_space(v.at(0))) {
        v.skip(1);
        if (v.size() >= 2 && v.at(0) == '/' && v.at(1) == '/') {
            while (v.size() && v.at(0) != '\n' && v.at(0) != '\r')
                v.skip(1);
            v.skip(1);
        }
    }
}

Params parse_block(View& block, int depth) {
    Params params;
    while (true) {
        escape_space(block);
        size_t keyBegin;
        if (!first_of(block, is_letter, keyBegin))
            break;
        size_t sep;
        if (!first_of(block, [](char c) { return c == ':' || c == '=' || c == '{'; }, sep))
            throw std::runtime_error("scene parse: expected ':' or '{' after a key");
        size_t blockEnd;
        if (first_of(block, [](char c) { return c == '}'; }, blockEnd) && sep > blockEnd)
            break;
        size_t keyEnd;
        if (!first_of(block, [](char c) { return !is_letter(c); }, keyEnd))
            throw std::runtime_error("scene parse: unterminated key");
        std::string key = block.str(0, keyEnd);
        std::string rawName = block.str(keyEnd, sep - keyEnd), name;
        for (char c : rawName)
            if (is_letter(c))
                name += c;
        block.skip(sep);
        bool typed = false;
        std::string value;
        if (block.at(0) != '{') {
            typed = true;
            block.skip(1);
            escape_space(block);
            size_t valueEnd;
            if (!first_of(block, [](char c) { return c == '\n' || c == '\r' || c == '{'; },
                          valueEnd))
                throw std::runtime_error("scene parse: value of \"" + key + "\" is not terminated");
            value = block.str(0, valueEnd);
            params.values[key] = value;
            block.skip(valueEnd);
        }
        escape_space(block);
        if (block.at(0) == '{') {
            block.skip(1);
            auto& list = params.subsets[key];
            list.push_back(parse_block(block, depth + 1));
            Params& sub = list.back();
            if (typed)
                sub.values["@"] = value;
            if (!name.empty() && !sub.has_value("name"))
                sub.values["name"] = name;
        }
    }
    size_t blockEnd;
    if (first_of(block, [](char c) { return c == '}'; }, blockEnd))
        block.skip(blockEnd + 1);
    return params;
}

}  // namespace

Params parse_scene_text(const std::string& text) {
    View v{&text, 0};
    return parse_block(v, 0);
}

}  // namespace pg
