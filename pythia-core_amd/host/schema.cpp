// schema.cpp -- plan configuration parsing, Schema, Page and the hash
// specification checks of the host operator layer.
#include <ctype.h>
#include <string.h>

#include <sstream>

#include "pjh.h"

namespace pjh {

// ---------------------------------------------------------------- Setting

bool Setting::exists(const std::string& n) const
{
    for (auto& c : children)
        if (c->name == n)
            return true;
    return false;
}

const Setting& Setting::operator[](const std::string& n) const
{
    for (auto& c : children)
        if (c->name == n)
            return *c;
    throw MissingParameterException(n + (name.empty() ? "" : " (in " + name + ")"));
}

const Setting& Setting::operator[](int i) const
{
    if (i < 0 || i >= (int)children.size())
        throw InvalidParameter("index " + std::to_string(i) + " of " + name);
    return *children[i];
}

int64_t Setting::asInt() const
{
    if (type == Int)
        return ival;
    if (type == Float)
        return (int64_t)fval;
    if (type == Bool)
        return ival;
    throw InvalidParameter(name + " is not a number");
}

std::string Setting::asString() const
{
    if (type != String)
        throw InvalidParameter(name + " is not a string");
    return sval;
}

namespace {

class Parser {
  public:
    explicit Parser(const std::string& t) : s(t) {}

    std::unique_ptr<Setting> parse()
    {
        auto root = std::unique_ptr<Setting>(new Setting());
        root->type = Setting::Group;
        for (;;) {
            skip();
            if (p >= s.size())
                break;
            root->children.push_back(setting());
        }
        return root;
    }

  private:
    const std::string& s;
    size_t p = 0;
    int line = 1;

    [[noreturn]] void error(const std::string& what)
    {
        throw PlanError("plan syntax error at line " + std::to_string(line) + ": " + what);
    }

    void skip()
    {
        while (p < s.size()) {
            char c = s[p];
            if (c == '\n') {
                line++;
                p++;
            } else if (isspace((unsigned char)c)) {
                p++;
            } else if (c == '#' || (c == '/' && p + 1 < s.size() && s[p + 1] == '/')) {
                while (p < s.size() && s[p] != '\n')
                    p++;
            } else if (c == '/' && p + 1 < s.size() && s[p + 1] == '*') {
                size_t e = s.find("*/", p + 2);
                if (e == std::string::npos)
                    error("unterminated comment");
                for (size_t i = p; i < e; ++i)
                    if (s[i] == '\n')
                        line++;
                p = e + 2;
            } else {
                break;
            }
        }
    }

    bool peek(char c)
    {
        skip();
        return p < s.size() && s[p] == c;
    }

    void expect(char c)
    {
        if (!peek(c))
            error(std::string("expected '") + c + "'");
        p++;
    }

    std::unique_ptr<Setting> setting()
    {
        skip();
        size_t b = p;
        if (p >= s.size() || !(isalpha((unsigned char)s[p]) || s[p] == '*'))
            error("expected a setting name");
        while (p < s.size() && (isalnum((unsigned char)s[p]) || s[p] == '_' || s[p] == '-' || s[p] == '*'))
            p++;
        std::string name = s.substr(b, p - b);
        skip();
        if (p < s.size() && (s[p] == '=' || s[p] == ':'))
            p++;
        else
            error("expected '=' or ':' after " + name);
        auto v = value();
        v->name = name;
        skip();
        if (p < s.size() && (s[p] == ';' || s[p] == ','))
            p++;
        return v;
    }

    std::unique_ptr<Setting> value()
    {
        skip();
        if (p >= s.size())
            error("expected a value");
        auto v = std::unique_ptr<Setting>(new Setting());
        char c = s[p];
        if (c == '{') {
            p++;
            v->type = Setting::Group;
            while (!peek('}')) {
                if (p >= s.size())
                    error("unterminated group");
                v->children.push_back(setting());
            }
            p++;
        } else if (c == '(' || c == '[') {
            const char close = c == '(' ? ')' : ']';
            p++;
            v->type = c == '(' ? Setting::List : Setting::Array;
            if (!peek(close)) {
                for (;;) {
                    auto e = value();
                    if (v->type == Setting::Array && e->isAggregate())
                        error("arrays hold scalars only");
                    v->children.push_back(std::move(e));
                    if (peek(',')) {
                        p++;
                        continue;
                    }
                    break;
                }
            }
            expect(close);
        } else if (c == '"') {
            v->type = Setting::String;
            while (peek('"')) {
                p++;
                while (p < s.size() && s[p] != '"') {
                    if (s[p] == '\\' && p + 1 < s.size()) {
                        char e = s[p + 1];
                        v->sval += e == 'n' ? '\n' : e == 't' ? '\t' : e;
                        p += 2;
                    } else {
                        v->sval += s[p++];
                    }
                }
                if (p >= s.size())
                    error("unterminated string");
                p++;
            }
        } else {
            size_t b = p;
            while (p < s.size() && (isalnum((unsigned char)s[p]) || s[p] == '.' || s[p] == '-' || s[p] == '+'))
                p++;
            std::string tok = s.substr(b, p - b);
            if (tok.empty())
                error("unexpected character");
            std::string low;
            for (char ch : tok)
                low += (char)tolower((unsigned char)ch);
            if (low == "true" || low == "false") {
                v->type = Setting::Bool;
                v->ival = low == "true";
            } else if (low.find_first_of(".e") != std::string::npos && low.compare(0, 2, "0x") != 0) {
                v->type = Setting::Float;
                v->fval = strtod(tok.c_str(), nullptr);
            } else {
                while (!low.empty() && low.back() == 'l')
                    low.pop_back();
                char* end = nullptr;
                v->type = Setting::Int;
                v->ival = (int64_t)strtoll(low.c_str(), &end, 0);
                if (!end || *end)
                    error("bad number " + tok);
            }
        }
        return v;
    }
};

}  // namespace

std::unique_ptr<Setting> parse_config(const std::string& text)
{
    return Parser(text).parse();
}

// ---------------------------------------------------------------- Schema

void Schema::add(ColumnType ct, unsigned width)
{
    unsigned w;
    switch (ct) {
    case CT_INTEGER: w = 4; break;
    case CT_LONG:
    case CT_DECIMAL:
    case CT_DATE:
    case CT_POINTER: w = 8; break;
    case CT_CHAR:
        if (width < 2)
            throw IllegalSchemaDeclarationException("char columns need a width of at least 2");
        w = width;
        break;
    default: throw IllegalSchemaDeclarationException("unknown column type");
    }
    types.push_back(ct);
    widths.push_back(w);
    offsets.push_back(size);
    size += w;
}

ColumnType Schema::parse(const std::string& s, unsigned* width)
{
    *width = 0;
    if (s == "int" || s == "integer")
        return CT_INTEGER;
    if (s == "long")
        return CT_LONG;
    if (s == "dec" || s == "decimal")
        return CT_DECIMAL;
    if (s == "date")
        return CT_DATE;
    if (s == "pointer")
        return CT_POINTER;
    if (s.compare(0, 5, "char(") == 0 && s.back() == ')') {
        *width = (unsigned)atoi(s.c_str() + 5);
        return CT_CHAR;
    }
    throw IllegalSchemaDeclarationException("unknown column type " + s);
}

void Schema::to_c(pj_schema* out) const
{
    memset(out, 0, sizeof(*out));
    if (types.size() > PJ_MAX_COLS)
        throw IllegalSchemaDeclarationException("too many columns");
    out->ncols = (uint32_t)types.size();
    for (size_t i = 0; i < types.size(); ++i) {
        out->type[i] = (uint32_t)types[i];
        out->width[i] = widths[i];
    }
}

// ---------------------------------------------------------------- Page

Page::Page(uint64_t bytes, unsigned tuplesize)
    : own(std::make_shared<std::vector<char>>(bytes)), cap(bytes), tsize(tuplesize)
{
    base = own->data();
    if (bytes < tuplesize)
        throw InvalidParameter("buffsize smaller than one tuple");
}

Page::Page(void* data, uint64_t bytes, unsigned tuplesize, bool full)
    : base(static_cast<char*>(data)), cap(bytes), used(full ? bytes : 0), tsize(tuplesize)
{
}

void* Page::allocateTuple()
{
    if (!canStoreTuple())
        return nullptr;
    void* r = base + used;
    used += tsize;
    return r;
}

// ---------------------------------------------------------------- hashing

unsigned getlogarithm(unsigned k)
{
    --k;
    for (unsigned i = 0; i < 32; ++i)
        if ((0xFFFFFFFFu << (31 - i)) & k)
            return 32 - i;
    return 0;
}

HashSpec make_hashspec(const Setting& node)
{
    HashSpec h;
    h.fn = node["fn"].asString();
    static const char* known[] = {"modulo", "parammodulo", "knuth", "range", "exactrange", "bytes",
                                  "tpchorderkey", "willis", "alwayszero", "tpchq1magic"};
    bool ok = false;
    for (const char* k : known)
        ok |= h.fn == k;
    if (!ok)
        throw UnknownHashException(h.fn);
    if (h.fn == "alwayszero") {
        h.buckets = h.rounded = 1;
        return h;
    }
    if (h.fn == "tpchq1magic") {
        h.buckets = h.rounded = 4;
        return h;
    }
    const int64_t b = node["buckets"].asInt();
    if (b <= 0)
        throw MissingParameterException("Number of hash buckets cannot be zero.");
    h.buckets = (unsigned)b;
    if (node.exists("offset"))
        h.offset = (unsigned)node["offset"].asInt();
    if (node.exists("skipbits"))
        h.skipbits = (unsigned)node["skipbits"].asInt();
    if (h.fn == "range" || h.fn == "exactrange") {
        const Setting& r = node["range"];
        h.range_min = r[0].asInt();
        h.range_max = r[1].asInt();
    }
    h.rounded = h.fn == "exactrange" ? h.buckets : 1u << getlogarithm(h.buckets);
    return h;
}

}  // namespace pjh
