// Recursive-descent reader for the XML subset described in xml.hpp.
#include "xml.hpp"

#include <cctype>
#include <fstream>
#include <sstream>

#include "common.hpp"

namespace etn {

const XmlNode* XmlNode::child(const std::string& n) const
{
    for (const auto& c : children)
        if (c->name == n) return c.get();
    return nullptr;
}

const std::string* XmlNode::attribute(const std::string& n) const
{
    for (const auto& a : attributes)
        if (a.first == n) return &a.second;
    return nullptr;
}

namespace {

class Reader {
public:
    Reader(const std::string& t, const std::string& src) : s_(t), src_(src) {}

    std::vector<std::unique_ptr<XmlNode>> document()
    {
        std::vector<std::unique_ptr<XmlNode>> out;
        if (s_.compare(0, 3, "\xEF\xBB\xBF") == 0) i_ = 3;   // UTF-8 byte order mark
        for (;;) {
            skip_misc();
            if (i_ >= s_.size()) break;
            if (s_[i_] != '<') {
                // top-level text (whitespace already skipped) is not allowed
                error("text outside an element");
            }
            out.push_back(element());
        }
        return out;
    }

private:
    [[noreturn]] void error(const std::string& what) const
    {
        size_t line = 1;
        for (size_t k = 0; k < i_ && k < s_.size(); ++k)
            if (s_[k] == '\n') ++line;
        fail(src_ + ":" + std::to_string(line) + ": malformed XML (" + what + ")");
    }

    bool starts(const char* p) const { return s_.compare(i_, std::char_traits<char>::length(p), p) == 0; }

    void skip_ws()
    {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
    }

    void skip_until(const char* end)
    {
        size_t k = s_.find(end, i_);
        if (k == std::string::npos) error(std::string("missing '") + end + "'");
        i_ = k + std::char_traits<char>::length(end);
    }

    // whitespace, comments, <?...?> and <!DOCTYPE ...>
    void skip_misc()
    {
        for (;;) {
            skip_ws();
            if (starts("<!--")) skip_until("-->");
            else if (starts("<?")) skip_until("?>");
            else if (starts("<!")) skip_until(">");
            else return;
        }
    }

    static bool name_char(char c)
    {
        return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == '.' || c == ':';
    }

    std::string name()
    {
        size_t b = i_;
        while (i_ < s_.size() && name_char(s_[i_])) ++i_;
        if (b == i_) error("expected a name");
        return s_.substr(b, i_ - b);
    }

    std::string unescape(const std::string& v)
    {
        std::string out;
        out.reserve(v.size());
        for (size_t k = 0; k < v.size(); ++k) {
            if (v[k] != '&') {
                out += v[k];
                continue;
            }
            size_t e = v.find(';', k);
            if (e == std::string::npos) error("unterminated entity");
            std::string ent = v.substr(k + 1, e - k - 1);
            if (ent == "amp") out += '&';
            else if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "quot") out += '"';
            else if (ent == "apos") out += '\'';
            else if (!ent.empty() && ent[0] == '#') {
                unsigned long cp = (ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X'))
                                       ? std::stoul(ent.substr(2), nullptr, 16)
                                       : std::stoul(ent.substr(1), nullptr, 10);
                if (cp < 0x80) out += (char)cp;
                else if (cp < 0x800) {
                    out += (char)(0xC0 | (cp >> 6));
                    out += (char)(0x80 | (cp & 0x3F));
                } else {
                    out += (char)(0xE0 | ((cp >> 12) & 0x0F));
                    out += (char)(0x80 | ((cp >> 6) & 0x3F));
                    out += (char)(0x80 | (cp & 0x3F));
                }
            } else error("unknown entity &" + ent + ";");
            k = e;
        }
        return out;
    }

    std::unique_ptr<XmlNode> element()
    {
        auto node = std::make_unique<XmlNode>();
        ++i_;   // '<'
        node->name = name();
        for (;;) {
            skip_ws();
            if (i_ >= s_.size()) error("unterminated start tag <" + node->name + ">");
            if (starts("/>")) {
                i_ += 2;
                return node;
            }
            if (s_[i_] == '>') {
                ++i_;
                break;
            }
            std::string an = name();
            skip_ws();
            if (i_ >= s_.size() || s_[i_] != '=') error("expected '=' after attribute " + an);
            ++i_;
            skip_ws();
            if (i_ >= s_.size() || (s_[i_] != '"' && s_[i_] != '\'')) error("expected a quoted value");
            char q = s_[i_++];
            size_t e = s_.find(q, i_);
            if (e == std::string::npos) error("unterminated attribute value");
            node->attributes.emplace_back(an, unescape(s_.substr(i_, e - i_)));
            i_ = e + 1;
        }
        // content
        for (;;) {
            if (i_ >= s_.size()) error("missing </" + node->name + ">");
            if (starts("<!--")) skip_until("-->");
            else if (starts("<![CDATA[")) skip_until("]]>");
            else if (starts("<?")) skip_until("?>");
            else if (starts("</")) {
                i_ += 2;
                std::string n = name();
                if (n != node->name) error("</" + n + "> closes <" + node->name + ">");
                skip_ws();
                if (i_ >= s_.size() || s_[i_] != '>') error("expected '>'");
                ++i_;
                return node;
            } else if (s_[i_] == '<') node->children.push_back(element());
            else ++i_;   // text content is not used
        }
    }

    const std::string& s_;
    std::string src_;
    size_t i_ = 0;
};

}  // namespace

std::vector<std::unique_ptr<XmlNode>> parse_xml(const std::string& text, const std::string& source_name)
{
    return Reader(text, source_name).document();
}

std::vector<std::unique_ptr<XmlNode>> parse_xml_file(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) fail("cannot open XML file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_xml(ss.str(), path);
}

std::string xml_escape(const std::string& s)
{
    std::string out;
    for (char c : s) {
        if (c == '&') out += "&amp;";
        else if (c == '<') out += "&lt;";
        else if (c == '>') out += "&gt;";
        else if (c == '"') out += "&quot;";
        else out += c;
    }
    return out;
}

}  // namespace etn
