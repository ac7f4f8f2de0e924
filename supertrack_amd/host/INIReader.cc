// INIReader.cc -- see INIReader.hh for the rules implemented here.
#include "INIReader.hh"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

namespace {

constexpr size_t kMaxLine = 200;     // characters per read, terminator included
constexpr size_t kMaxSection = 50;
constexpr size_t kMaxName = 50;

bool is_space(char c) { return std::isspace((unsigned char)c) != 0; }

std::string rstrip(const std::string &s)
{
    size_t e = s.size();
    while (e > 0 && is_space(s[e - 1])) --e;
    return s.substr(0, e);
}

size_t lskip(const std::string &s, size_t i)
{
    while (i < s.size() && is_space(s[i])) ++i;
    return i;
}

// Position of the first character of `chars` or of an inline comment (';'
// after white space) from i on; s.size() when neither occurs.
size_t find_chars_or_comment(const std::string &s, size_t i, const char *chars)
{
    bool was_space = false;
    for (; i < s.size(); ++i) {
        const char c = s[i];
        if (chars && std::strchr(chars, c)) break;
        if (was_space && c == ';') break;
        was_space = is_space(c);
    }
    return i;
}

// Splits text the way fgets with a kMaxLine buffer reads a stream.
std::vector<std::string> read_pieces(const std::string &text)
{
    std::vector<std::string> out;
    size_t i = 0;
    while (i < text.size()) {
        size_t n = 0;
        std::string piece;
        while (i < text.size() && n < kMaxLine - 1) {
            const char c = text[i++];
            piece.push_back(c);
            ++n;
            if (c == '\n') break;
        }
        out.push_back(piece);
    }
    return out;
}

}  // namespace

INIReader::INIReader(const std::string &filename)
{
    std::ifstream f(filename, std::ios::binary);
    if (!f) {
        _error = -1;
        return;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    Parse(ss.str());
}

INIReader INIReader::FromString(const std::string &text)
{
    INIReader r;
    r.Parse(text);
    return r;
}

void INIReader::Parse(const std::string &text)
{
    std::string section, prev_name;
    int lineno = 0;
    for (const std::string &raw : read_pieces(text)) {
        ++lineno;
        std::string line = raw;
        if (lineno == 1 && line.size() >= 3 && (unsigned char)line[0] == 0xEF && (unsigned char)line[1] == 0xBB &&
            (unsigned char)line[2] == 0xBF)
            line = line.substr(3);
        line = rstrip(line);
        const size_t start = lskip(line, 0);
        if (start >= line.size()) continue;
        const char c0 = line[start];
        if (c0 == ';' || c0 == '#') continue;
        if (!prev_name.empty() && start > 0) {
            // continuation of the previous value
            size_t end = find_chars_or_comment(line, start, nullptr);
            Store(section, prev_name, rstrip(line.substr(start, end - start)));
            continue;
        }
        if (c0 == '[') {
            size_t end = find_chars_or_comment(line, start + 1, "]");
            if (end < line.size() && line[end] == ']') {
                section = line.substr(start + 1, end - start - 1).substr(0, kMaxSection - 1);
                prev_name.clear();
            } else if (!_error) {
                _error = lineno;
            }
            continue;
        }
        size_t end = find_chars_or_comment(line, start, "=:");
        if (end < line.size() && (line[end] == '=' || line[end] == ':')) {
            const std::string name = rstrip(line.substr(start, end - start));
            size_t v0 = lskip(line, end + 1);
            size_t v1 = find_chars_or_comment(line, v0, nullptr);
            const std::string value = rstrip(line.substr(v0, v1 - v0));
            prev_name = name.substr(0, kMaxName - 1);
            Store(section, name, value);
        } else if (!_error) {
            _error = lineno;
        }
    }
}

std::string INIReader::MakeKey(const std::string &section, const std::string &name)
{
    std::string key = section + "=" + name;
    std::transform(key.begin(), key.end(), key.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return key;
}

void INIReader::Store(const std::string &section, const std::string &name, const std::string &value)
{
    std::string &v = _values[MakeKey(section, name)];
    if (!v.empty()) v += "\n";
    v += value;
    _sections.insert(section);
}

std::string INIReader::Get(std::string section, std::string name, std::string default_value) const
{
    auto it = _values.find(MakeKey(section, name));
    return it == _values.end() ? default_value : it->second;
}

long INIReader::GetInteger(std::string section, std::string name, long default_value) const
{
    const std::string s = Get(section, name, "");
    char *end = nullptr;
    const long n = std::strtol(s.c_str(), &end, 0);
    return end > s.c_str() ? n : default_value;
}

double INIReader::GetReal(std::string section, std::string name, double default_value) const
{
    const std::string s = Get(section, name, "");
    char *end = nullptr;
    const double n = std::strtod(s.c_str(), &end);
    return end > s.c_str() ? n : default_value;
}

float INIReader::GetFloat(std::string section, std::string name, float default_value) const
{
    const std::string s = Get(section, name, "");
    char *end = nullptr;
    const float n = std::strtof(s.c_str(), &end);
    return end > s.c_str() ? n : default_value;
}

bool INIReader::GetBoolean(std::string section, std::string name, bool default_value) const
{
    std::string s = Get(section, name, "");
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    if (s == "true" || s == "yes" || s == "on" || s == "1") return true;
    if (s == "false" || s == "no" || s == "off" || s == "0") return false;
    return default_value;
}
