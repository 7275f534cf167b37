// json.cpp -- small JSON reader for the experiment configs (replaces the
// nlohmann::json parse the reference drivers use, e.g. FullExperiment.cpp:143-144).
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "twisty_host.hpp"

namespace twisty {

class JsonParser {
   public:
    explicit JsonParser(const std::string &t) : s(t) { }

    Json parse()
    {
        Json v = value();
        ws();
        if (i != s.size()) fail("trailing characters");
        return v;
    }

   private:
    const std::string &s;
    size_t i = 0;

    [[noreturn]] void fail(const std::string &what) const
    {
        throw std::runtime_error("JSON parse error at offset " + std::to_string(i) + ": " + what);
    }
    void ws()
    {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool lit(const char *w)
    {
        size_t n = 0;
        while (w[n]) ++n;
        if (s.compare(i, n, w) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    Json value()
    {
        ws();
        if (i >= s.size()) fail("unexpected end");
        Json v;
        const char c = s[i];
        if (c == '{') {
            v.m_type = Json::Type::Object;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') {
                ++i;
                return v;
            }
            for (;;) {
                ws();
                if (i >= s.size() || s[i] != '"') fail("expected key");
                std::string k = str();
                ws();
                if (i >= s.size() || s[i] != ':') fail("expected ':'");
                ++i;
                v.m_object[k] = value();
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    ++i;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.m_type = Json::Type::Array;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') {
                ++i;
                return v;
            }
            for (;;) {
                v.m_array.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',') {
                    ++i;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    ++i;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.m_type = Json::Type::String;
            v.m_string = str();
            return v;
        }
        if (lit("true")) {
            v.m_type = Json::Type::Bool;
            v.m_bool = true;
            return v;
        }
        if (lit("false")) {
            v.m_type = Json::Type::Bool;
            return v;
        }
        if (lit("null")) return v;
        const char *b = s.c_str() + i;
        char *e = nullptr;
        const double d = std::strtod(b, &e);
        if (e == b) fail("unexpected character");
        i += (size_t)(e - b);
        v.m_type = Json::Type::Number;
        v.m_number = d;
        return v;
    }
    std::string str()
    {
        ++i;  // opening quote
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) fail("bad escape");
                const char e = s[i++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (i + 4 > s.size()) fail("bad \\u escape");
                        const unsigned cp = (unsigned)std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
                        i += 4;
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
};

Json Json::Parse(const std::string &text) { return JsonParser(text).parse(); }

Json Json::ParseFile(const std::string &path)
{
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Failed to open: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return Parse(ss.str());
}

bool Json::contains(const std::string &key) const
{
    return m_type == Type::Object && m_object.count(key) != 0;
}

const Json &Json::at(const std::string &key) const
{
    if (m_type != Type::Object) throw std::runtime_error("JSON: '" + key + "' looked up in a non-object");
    auto it = m_object.find(key);
    if (it == m_object.end()) throw std::runtime_error("JSON: missing key '" + key + "'");
    return it->second;
}

const Json &Json::at(size_t idx) const
{
    if (m_type != Type::Array || idx >= m_array.size()) throw std::runtime_error("JSON: array index out of range");
    return m_array[idx];
}

size_t Json::size() const
{
    if (m_type == Type::Array) return m_array.size();
    if (m_type == Type::Object) return m_object.size();
    return 0;
}

double Json::number() const
{
    if (m_type == Type::Number) return m_number;
    if (m_type == Type::Bool) return m_bool ? 1.0 : 0.0;
    throw std::runtime_error("JSON: expected a number");
}

bool Json::boolean() const
{
    if (m_type == Type::Bool) return m_bool;
    if (m_type == Type::Number) return m_number != 0.0;
    throw std::runtime_error("JSON: expected a boolean");
}

const std::string &Json::string() const
{
    if (m_type != Type::String) throw std::runtime_error("JSON: expected a string");
    return m_string;
}

}  // namespace twisty
