#include "SceneParser.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace oxr_json {
namespace {

struct Reader {
    const std::string& s;
    size_t i = 0;
    [[noreturn]] void fail(const char* what) {
        throw std::runtime_error(std::string("json: ") + what + " at byte " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) { i++; return true; }
        return false;
    }
    Value value(int depth) {
        if (depth > 64) fail("nesting too deep");
        ws();
        if (i >= s.size()) fail("unexpected end");
        char c = s[i];
        Value v;
        if (c == '{') {
            i++;
            v.kind = Value::Object;
            if (eat('}')) return v;
            do {
                ws();
                if (i >= s.size() || s[i] != '"') fail("expected key");
                std::string k = string();
                if (!eat(':')) fail("expected ':'");
                v.obj[k] = value(depth + 1);
            } while (eat(','));
            if (!eat('}')) fail("expected '}'");
        } else if (c == '[') {
            i++;
            v.kind = Value::Array;
            if (eat(']')) return v;
            do v.arr.push_back(value(depth + 1));
            while (eat(','));
            if (!eat(']')) fail("expected ']'");
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string();
        } else if (s.compare(i, 4, "true") == 0) {
            v.kind = Value::Bool; v.b = true; i += 4;
        } else if (s.compare(i, 5, "false") == 0) {
            v.kind = Value::Bool; i += 5;
        } else if (s.compare(i, 4, "null") == 0) {
            i += 4;
        } else {
            const char* b = s.c_str() + i;
            char* e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) fail("unexpected character");
            v.kind = Value::Number;
            i += (size_t)(e - b);
        }
        return v;
    }
    std::string string() {
        i++;  // opening quote
        std::string out;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\') {
                if (++i >= s.size()) fail("bad escape");
                char e = s[i];
                out += e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e == 'b' ? '\b' : e == 'f' ? '\f' : e;
                if (e == 'u') fail("\\u escapes are not supported");
                i++;
            } else {
                out += s[i++];
            }
        }
        if (i >= s.size()) fail("unterminated string");
        i++;
        return out;
    }
};

}  // namespace

Value parse(const std::string& text) {
    Reader r{text};
    Value v = r.value(0);
    r.ws();
    if (r.i != text.size()) r.fail("trailing characters");
    return v;
}

}  // namespace oxr_json

void SceneParser::LoadFromJson(const std::string& fn, Scene& scene, bool& is_succeed) {
    is_succeed = false;
    std::ifstream in(fn);
    if (!in.is_open()) return;  // SceneParser.h:31-35: missing file -> is_succeed = false
    std::stringstream buf;
    buf << in.rdbuf();
    oxr_json::Value data = oxr_json::parse(buf.str());
    is_succeed = true;
    Parse(data, scene);
}

void SceneParser::Parse(const oxr_json::Value& data, Scene&) {
    width = (int)data.number("width", 1024);
    height = (int)data.number("height", 1024);
    depth = (int)data.number("depth", 4);
    const oxr_json::Value* f = data.get("filter");
    oxr_json::Value empty;
    ParseFilter(f ? *f : empty);
}

void SceneParser::ParseFilter(const oxr_json::Value& data) {
    std::string t = data.string("type", "gaussian");
    if (t == "triangle") filterType = FilterType::Triangle;
    else if (t == "tent") filterType = FilterType::Tent;
    else if (t == "box") filterType = FilterType::Box;
    else filterType = FilterType::Gaussian;
}
