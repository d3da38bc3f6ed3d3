// The reference's XML configuration file (CConfiguration.cpp:54-197) for the
// native driver: a small XML element reader (elements, text, the five
// predefined entities; declarations and comments skipped) and the
// configuration it describes. Numbers go through the C library's atof / atoi
// and are narrowed to the arithmetic type T, as the reference stores them;
// `domian-length` keeps the reference's spelling; `subdomain-num` is read as
// in the reference's MPI build. check() applies checkParameters
// (CConfiguration.cpp:199-222) and throws where the reference asserts.
#pragma once

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace lbm {

struct XmlNode {
    std::string name;
    std::string text;  // character data before the first child element
    std::vector<std::unique_ptr<XmlNode>> children;

    const XmlNode* child(const std::string& n) const
    {
        for (const auto& c : children)
            if (c->name == n)
                return c.get();
        return nullptr;
    }
};

class XmlReader {
public:
    explicit XmlReader(const std::string& s) : s_(s) {}

    std::unique_ptr<XmlNode> document()
    {
        skip_misc();
        auto root = element();
        skip_misc();
        if (i_ != s_.size())
            fail("content after the root element");
        return root;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const std::string& what) const
    {
        throw std::runtime_error("XML parse error at byte " + std::to_string(i_) + ": " + what);
    }
    bool starts(const char* p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
    void skip_ws()
    {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_]))
            i_++;
    }
    void skip_until(const char* end)
    {
        const size_t k = s_.find(end, i_);
        if (k == std::string::npos)
            fail(std::string("missing ") + end);
        i_ = k + std::strlen(end);
    }
    // whitespace, <?...?>, <!-- ... -->, <!DOCTYPE ...>
    void skip_misc()
    {
        for (;;) {
            skip_ws();
            if (starts("<?"))
                skip_until("?>");
            else if (starts("<!--"))
                skip_until("-->");
            else if (starts("<!"))
                skip_until(">");
            else
                return;
        }
    }
    std::string name()
    {
        const size_t b = i_;
        while (i_ < s_.size() && !std::isspace((unsigned char)s_[i_]) && s_[i_] != '>' &&
               s_[i_] != '/' && s_[i_] != '=')
            i_++;
        if (b == i_)
            fail("expected a name");
        return s_.substr(b, i_ - b);
    }
    static std::string decode(const std::string& t)
    {
        static const char* ent[][2] = {
            {"&lt;", "<"}, {"&gt;", ">"}, {"&amp;", "&"}, {"&quot;", "\""}, {"&apos;", "'"}};
        std::string out;
        for (size_t k = 0; k < t.size();) {
            bool hit = false;
            if (t[k] == '&')
                for (auto& e : ent)
                    if (t.compare(k, std::strlen(e[0]), e[0]) == 0) {
                        out += e[1];
                        k += std::strlen(e[0]);
                        hit = true;
                        break;
                    }
            if (!hit)
                out += t[k++];
        }
        return out;
    }
    std::unique_ptr<XmlNode> element()
    {
        if (i_ >= s_.size() || s_[i_] != '<')
            fail("expected '<'");
        i_++;
        auto node = std::make_unique<XmlNode>();
        node->name = name();
        // attributes are not used by the configuration format: skip them
        for (;;) {
            skip_ws();
            if (i_ >= s_.size())
                fail("unterminated tag");
            if (starts("/>")) {
                i_ += 2;
                return node;
            }
            if (s_[i_] == '>') {
                i_++;
                break;
            }
            name();
            skip_ws();
            if (i_ >= s_.size() || s_[i_] != '=')
                fail("expected '='");
            i_++;
            skip_ws();
            const char q = i_ < s_.size() ? s_[i_] : 0;
            if (q != '"' && q != '\'')
                fail("expected a quoted value");
            const size_t e = s_.find(q, i_ + 1);
            if (e == std::string::npos)
                fail("unterminated attribute value");
            i_ = e + 1;
        }
        bool text_done = false;
        for (;;) {
            const size_t lt = s_.find('<', i_);
            if (lt == std::string::npos)
                fail("unterminated element <" + node->name + ">");
            if (!text_done)
                node->text += decode(s_.substr(i_, lt - i_));
            i_ = lt;
            if (starts("</")) {
                i_ += 2;
                if (name() != node->name)
                    fail("mismatched end tag for <" + node->name + ">");
                skip_ws();
                if (i_ >= s_.size() || s_[i_] != '>')
                    fail("expected '>'");
                i_++;
                return node;
            }
            if (starts("<!--")) {
                skip_until("-->");
                continue;
            }
            if (starts("<?")) {
                skip_until("?>");
                continue;
            }
            node->children.push_back(element());
            text_done = true;
        }
    }
};

struct Configuration {
    double velocity[3] = {0, 0, 0}, acceleration[3] = {0, 0, 0};
    double viscosity = -1, max_velocity = 0.1, max_acceleration = 0.1;
    int domain_size[3] = {32, 32, 32};
    double domain_length[3] = {32, 32, 32};
    int subdomains[3] = {1, 1, 1};
    double cpu_subdomain_ratio = 0;
    int loops = 100;
    double timestep = 1;
    int do_benchmark = 0, do_logging = 0, do_validation = 0, do_visualization = 0;
    std::string benchmark_dir, logging_dir, validation_dir, visualization_dir;
    int visualization_rate = 1;
    int cpu_blocks[2][3] = {{32, 32, 32}, {32, 32, 32}};
    int gpu_grids[3][3] = {{32, 16, 1}, {32, 8, 1}, {32, 8, 1}};
    bool f64 = false;

    bool store_outputs() const { return do_validation || do_visualization; }
    double narrow(double v) const { return f64 ? v : (double)(float)v; }

    // size, origin (cells) and length of one rank's subdomain, x fastest
    // (CManager.cpp:34-58)
    void subdomain(int rank, int size[3], int origin[3], double length[3]) const
    {
        const int c[3] = {rank % subdomains[0], (rank / subdomains[0]) % subdomains[1],
                          rank / (subdomains[0] * subdomains[1])};
        for (int a = 0; a < 3; a++) {
            size[a] = domain_size[a] / subdomains[a];
            origin[a] = c[a] * size[a];
            length[a] = f64 ? domain_length[a] / (double)subdomains[a]
                            : (double)((float)domain_length[a] / (float)subdomains[a]);
        }
    }

    void check() const
    {
        auto need = [](bool ok, const char* what) {
            if (!ok)
                throw std::runtime_error(std::string("configuration check failed: ") + what);
        };
        need(timestep > 0, "timestep > 0");
        need(max_velocity > 0, "max velocity > 0");
        need(max_acceleration > 0, "max acceleration > 0");
        for (int a = 0; a < 3; a++) {
            need(domain_size[a] > 0, "domain size > 0");
            need(domain_length[a] > 0, "domain length > 0");
            need(subdomains[a] > 0, "subdomain count > 0");
            need(domain_size[a] % subdomains[a] == 0, "domain size divisible by the subdomain count");
        }
        need(cpu_subdomain_ratio >= 0 && cpu_subdomain_ratio <= 1, "0 <= CPU subdomain ratio <= 1");
        for (int a = 0; a < 3; a++) {
            const int b = (a + 1) % 3;
            const double d = f64 ? std::fabs(domain_length[a] / domain_size[a] -
                                             domain_length[b] / domain_size[b])
                                 : std::fabs((float)domain_length[a] / (float)domain_size[a] -
                                             (float)domain_length[b] / (float)domain_size[b]);
            const double eps = f64 ? std::numeric_limits<double>::epsilon()
                                   : std::numeric_limits<float>::epsilon();
            need(d < eps, "cubic cells (equal length / size on every axis)");
        }
        need(visualization_rate > 0, "visualization rate > 0");
        for (auto& b : cpu_blocks)
            need(b[0] > 0 && b[1] > 0 && b[2] > 0, "CPU block sizes > 0");
        for (auto& g : gpu_grids) {
            need(g[0] > 0 && g[1] > 0 && g[2] > 0, "GPU grid sizes > 0");
            need(g[0] * g[1] * g[2] <= 1024, "GPU grid <= 1024 threads");
        }
    }
};

inline Configuration load_configuration(const std::string& path, bool f64)
{
    std::ifstream in(path, std::ios::binary);
    if (!in)
        throw std::runtime_error("Loading XML configuration file \"" + path + "\" failed");
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    XmlReader r(text);
    auto root = r.document();
    if (root->name != "lbm-configuration")
        throw std::runtime_error(path + ": root element is not <lbm-configuration>");
    auto node = [&](std::initializer_list<const char*> names) {
        const XmlNode* n = root.get();
        std::string where;
        for (const char* s : names) {
            where += (where.empty() ? "" : "/") + std::string(s);
            n = n->child(s);
            if (!n)
                throw std::runtime_error(path + ": missing element " + where);
        }
        return n;
    };
    Configuration c;
    c.f64 = f64;
    auto f = [&](std::initializer_list<const char*> names) {
        return c.narrow(std::atof(node(names)->text.c_str()));
    };
    auto i = [&](std::initializer_list<const char*> names) {
        return std::atoi(node(names)->text.c_str());
    };
    auto s = [&](std::initializer_list<const char*> names) { return node(names)->text; };
    const char* xyz[3] = {"x", "y", "z"};
    for (int a = 0; a < 3; a++) {
        c.velocity[a] = f({"physics", "velocity", xyz[a]});
        c.acceleration[a] = f({"physics", "acceleration", xyz[a]});
        c.domain_size[a] = i({"grid", "domain-size", xyz[a]});
        c.domain_length[a] = f({"grid", "domian-length", xyz[a]});
        c.subdomains[a] = i({"grid", "subdomain-num", xyz[a]});
        c.cpu_blocks[0][a] = i({"cpu", "block-configuration", "alpha-block-configuration", xyz[a]});
        c.cpu_blocks[1][a] = i({"cpu", "block-configuration", "beta-block-configuration", xyz[a]});
        c.gpu_grids[0][a] = i({"gpu", "grid-configuration", "init-grid-configuration", xyz[a]});
        c.gpu_grids[1][a] = i({"gpu", "grid-configuration", "alpha-grid-configuration", xyz[a]});
        c.gpu_grids[2][a] = i({"gpu", "grid-configuration", "beta-grid-configuration", xyz[a]});
    }
    c.viscosity = f({"physics", "viscosity"});
    c.max_velocity = f({"physics", "max-velocity"});
    c.max_acceleration = f({"physics", "max-acceleration"});
    c.cpu_subdomain_ratio = f({"grid", "cpu-subdomain-ratio"});
    c.loops = i({"simulation", "loops"});
    c.timestep = f({"simulation", "timestep"});
    c.do_benchmark = i({"simulation", "benchmark", "do"});
    c.do_logging = i({"simulation", "logging", "do"});
    c.do_validation = i({"simulation", "validation", "do"});
    c.do_visualization = i({"simulation", "visualization", "do"});
    c.benchmark_dir = s({"simulation", "benchmark", "output-dir"});
    c.logging_dir = s({"simulation", "logging", "output-dir"});
    c.validation_dir = s({"simulation", "validation", "output-dir"});
    c.visualization_dir = s({"simulation", "visualization", "output-dir"});
    c.visualization_rate = i({"simulation", "visualization", "rate"});
    c.check();
    return c;
}

}  // namespace lbm
