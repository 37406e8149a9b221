// xml_reader.cpp -- see xml_reader.h.
#include "xml_reader.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace nexusmd {

const std::string* XmlAttrs::find(const char* name) const {
    for (const auto& kv : items)
        if (kv.first == name) return &kv.second;
    return nullptr;
}

double XmlAttrs::as_double(const char* name) const {
    const std::string* v = find(name);
    if (!v) return 0.0;
    const char* s = v->c_str();
    while (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r') ++s;
    char* end = nullptr;
    const double d = std::strtod(s, &end);
    return end == s ? 0.0 : d;
}

int XmlAttrs::as_int(const char* name) const {
    const std::string* v = find(name);
    if (!v) return 0;
    const char* s = v->c_str();
    while (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r') ++s;
    char* end = nullptr;
    const long l = std::strtol(s, &end, 10);
    return end == s ? 0 : static_cast<int>(l);
}

std::string XmlAttrs::as_string(const char* name) const {
    const std::string* v = find(name);
    return v ? *v : std::string();
}

namespace {

struct Scanner {
    const char* p;
    const char* end;
    const char* begin;

    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("XML parse error at byte ") + std::to_string(p - begin) + ": " + what);
    }
    bool at(const char* lit) const {
        const size_t n = std::strlen(lit);
        return static_cast<size_t>(end - p) >= n && std::memcmp(p, lit, n) == 0;
    }
    void skip_past(const char* lit) {
        const size_t n = std::strlen(lit);
        while (static_cast<size_t>(end - p) >= n) {
            if (std::memcmp(p, lit, n) == 0) { p += n; return; }
            ++p;
        }
        fail("unterminated construct");
    }
    static bool space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }
    static bool name_char(char c) { return !space(c) && c != '=' && c != '>' && c != '/' && c != '<' && c != '"' && c != '\''; }
    void skip_space() { while (p < end && space(*p)) ++p; }
    std::string name() {
        const char* s = p;
        while (p < end && name_char(*p)) ++p;
        if (p == s) fail("expected a name");
        return std::string(s, p);
    }
};

void append_utf8(std::string& out, unsigned long cp) {
    if (cp < 0x80) out += static_cast<char>(cp);
    else if (cp < 0x800) { out += static_cast<char>(0xC0 | (cp >> 6)); out += static_cast<char>(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) {
        out += static_cast<char>(0xE0 | (cp >> 12));
        out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
        out += static_cast<char>(0x80 | (cp & 0x3F));
    } else {
        out += static_cast<char>(0xF0 | (cp >> 18));
        out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
        out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
        out += static_cast<char>(0x80 | (cp & 0x3F));
    }
}

std::string decode(const char* s, const char* e) {
    std::string out;
    out.reserve(static_cast<size_t>(e - s));
    while (s < e) {
        if (*s != '&') { out += *s++; continue; }
        const char* semi = static_cast<const char*>(std::memchr(s, ';', static_cast<size_t>(e - s)));
        if (!semi) { out += *s++; continue; }
        const std::string ent(s + 1, semi);
        if (ent == "lt") out += '<';
        else if (ent == "gt") out += '>';
        else if (ent == "amp") out += '&';
        else if (ent == "quot") out += '"';
        else if (ent == "apos") out += '\'';
        else if (!ent.empty() && ent[0] == '#') {
            const bool hex = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X');
            append_utf8(out, std::strtoul(ent.c_str() + (hex ? 2 : 1), nullptr, hex ? 16 : 10));
        } else {
            out.append(s, semi + 1);
        }
        s = semi + 1;
    }
    return out;
}

}  // namespace

void xml_scan(const char* data, size_t size, const XmlStartFn& on_start) {
    Scanner sc{data, data + size, data};
    std::vector<std::string> path;
    XmlAttrs attrs;
    // a UTF-8 byte-order mark is allowed before the prolog
    if (size >= 3 && static_cast<unsigned char>(data[0]) == 0xEF && static_cast<unsigned char>(data[1]) == 0xBB &&
        static_cast<unsigned char>(data[2]) == 0xBF)
        sc.p += 3;
    while (sc.p < sc.end) {
        const char* lt = static_cast<const char*>(std::memchr(sc.p, '<', static_cast<size_t>(sc.end - sc.p)));
        if (!lt) break;                       // trailing text
        sc.p = lt;
        if (sc.at("<!--")) { sc.skip_past("-->"); continue; }
        if (sc.at("<?")) { sc.skip_past("?>"); continue; }
        if (sc.at("<![CDATA[")) { sc.skip_past("]]>"); continue; }
        if (sc.at("<!")) { sc.skip_past(">"); continue; }
        if (sc.at("</")) {
            sc.p += 2;
            const std::string n = sc.name();
            sc.skip_space();
            if (sc.p >= sc.end || *sc.p != '>') sc.fail("expected '>' after a closing tag");
            ++sc.p;
            if (path.empty() || path.back() != n) sc.fail(("mismatched closing tag </" + n + ">").c_str());
            path.pop_back();
            continue;
        }
        ++sc.p;                               // '<'
        std::string n = sc.name();
        attrs.items.clear();
        bool self_closing = false;
        for (;;) {
            sc.skip_space();
            if (sc.p >= sc.end) sc.fail("unterminated start tag");
            if (*sc.p == '>') { ++sc.p; break; }
            if (*sc.p == '/') {
                ++sc.p;
                if (sc.p >= sc.end || *sc.p != '>') sc.fail("expected '>' after '/'");
                ++sc.p;
                self_closing = true;
                break;
            }
            std::string an = sc.name();
            sc.skip_space();
            if (sc.p >= sc.end || *sc.p != '=') sc.fail("expected '=' after an attribute name");
            ++sc.p;
            sc.skip_space();
            if (sc.p >= sc.end || (*sc.p != '"' && *sc.p != '\'')) sc.fail("expected a quoted attribute value");
            const char q = *sc.p++;
            const char* vs = sc.p;
            const char* ve = static_cast<const char*>(std::memchr(vs, q, static_cast<size_t>(sc.end - vs)));
            if (!ve) sc.fail("unterminated attribute value");
            attrs.items.emplace_back(std::move(an), decode(vs, ve));
            sc.p = ve + 1;
        }
        path.push_back(std::move(n));
        on_start(path, attrs);
        if (self_closing) path.pop_back();
    }
    if (!path.empty()) throw std::runtime_error("XML parse error: element <" + path.back() + "> is not closed");
}

bool xml_scan_file(const std::string& filename, const XmlStartFn& on_start, std::string* error) {
    FILE* f = std::fopen(filename.c_str(), "rb");
    if (!f) {
        if (error) *error = "cannot open " + filename + ": " + std::strerror(errno);
        return false;
    }
    std::string buf;
    if (std::fseek(f, 0, SEEK_END) == 0) {
        const long n = std::ftell(f);
        if (n > 0) buf.resize(static_cast<size_t>(n));
        std::fseek(f, 0, SEEK_SET);
    }
    const size_t got = buf.empty() ? 0 : std::fread(&buf[0], 1, buf.size(), f);
    std::fclose(f);
    buf.resize(got);
    xml_scan(buf.data(), buf.size(), on_start);
    return true;
}

}  // namespace nexusmd
