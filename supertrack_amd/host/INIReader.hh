// INIReader.hh -- .ini macro reader with the parsing rules SuperTrack's macros
// rely on (the vendored inih: extern/inih/INIReader.h:120-275, 395-460):
//   * lines are read in pieces of at most 199 characters; each is stripped of
//     surrounding white space;
//   * a line starting with ';' or '#' is a comment;
//   * "[name]" opens a section (a missing ']' is an error on that line);
//   * "name = value" or "name : value"; an inline comment starts at a ';'
//     that follows white space; values and names are right-stripped;
//   * a non-blank line that starts with white space continues the previous
//     name's value;
//   * section and name lookups ignore case; a repeated key appends "\n" and
//     the new value;
//   * GetInteger uses strtol base 0, GetReal strtod, GetFloat strtof, each
//     falling back to the default when nothing parses; GetBoolean accepts
//     true/yes/on/1 and false/no/off/0.
#pragma once

#include <map>
#include <set>
#include <string>

class INIReader {
public:
    INIReader() = default;
    explicit INIReader(const std::string &filename);
    // Parses text held in memory (same rules as a file).
    static INIReader FromString(const std::string &text);

    // 0 on success, -1 if the file could not be opened, else the first line
    // with an error.
    int ParseError() const { return _error; }
    const std::set<std::string> &Sections() const { return _sections; }

    std::string Get(std::string section, std::string name, std::string default_value) const;
    long GetInteger(std::string section, std::string name, long default_value) const;
    double GetReal(std::string section, std::string name, double default_value) const;
    float GetFloat(std::string section, std::string name, float default_value) const;
    bool GetBoolean(std::string section, std::string name, bool default_value) const;

private:
    static std::string MakeKey(const std::string &section, const std::string &name);
    void Parse(const std::string &text);
    void Store(const std::string &section, const std::string &name, const std::string &value);

    int _error = 0;
    std::map<std::string, std::string> _values;
    std::set<std::string> _sections;
};
