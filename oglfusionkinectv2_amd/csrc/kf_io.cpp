// kf_io.cpp -- input adapters around the frame loop (host code only):
// the calibration loaders of kinect.cpp:60-161 (OpenCV-style XML with
// Camera_Matrix / Distortion_Coefficients, or JSON with camera_matrix /
// dist_coeff) without the XML / JSON libraries the reference links.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kfusion.h"

int kf_fail(int code, const char *msg);

namespace {

bool read_file(const char *path, std::string &out)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, n);
    std::fclose(f);
    return true;
}

// splitStringToVector (kinect.cpp:60-76): commas removed, then every
// whitespace-separated number in order
std::vector<float> split_numbers(const std::string &text)
{
    std::string s;
    for (char c : text)
        if (c != ',') s.push_back(c);
    std::vector<float> out;
    const char *p = s.c_str();
    for (;;) {
        while (*p && std::isspace((unsigned char)*p)) ++p;
        if (!*p) break;
        char *end = nullptr;
        const float v = std::strtof(p, &end);
        if (end == p) break;  // the istringstream stops at the first non-number
        out.push_back(v);
        p = end;
    }
    return out;
}

// text of the first <data> element inside the first <name ...> element
bool xml_data(const std::string &doc, const char *name, std::string &out)
{
    const std::string open = std::string("<") + name;
    size_t a = doc.find(open);
    while (a != std::string::npos) {
        const char c = doc[a + open.size()];
        if (c == '>' || std::isspace((unsigned char)c) || c == '/') break;
        a = doc.find(open, a + 1);
    }
    if (a == std::string::npos) return false;
    const std::string close = std::string("</") + name + ">";
    const size_t b = doc.find(close, a);
    const std::string el = doc.substr(a, b == std::string::npos ? std::string::npos : b - a);
    const size_t d0 = el.find("<data>");
    if (d0 == std::string::npos) return false;
    const size_t d1 = el.find("</data>", d0);
    if (d1 == std::string::npos) return false;
    out = el.substr(d0 + 6, d1 - d0 - 6);
    return true;
}

// the numbers of the JSON value after "key": (nested arrays flattened in order)
bool json_numbers(const std::string &doc, const char *key, std::vector<double> &out)
{
    const std::string k = std::string("\"") + key + "\"";
    size_t a = doc.find(k);
    if (a == std::string::npos) return false;
    a = doc.find(':', a + k.size());
    if (a == std::string::npos) return false;
    size_t i = a + 1;
    while (i < doc.size() && std::isspace((unsigned char)doc[i])) ++i;
    if (i >= doc.size() || doc[i] != '[') return false;
    int depth = 0;
    for (; i < doc.size(); ++i) {
        const char c = doc[i];
        if (c == '[') {
            ++depth;
        } else if (c == ']') {
            if (--depth == 0) return true;
        } else if (c == '-' || c == '+' || c == '.' || std::isdigit((unsigned char)c)) {
            char *end = nullptr;
            const double v = std::strtod(doc.c_str() + i, &end);
            if (end == doc.c_str() + i) return false;
            out.push_back(v);
            i = (size_t)(end - doc.c_str()) - 1;
        }
    }
    return false;
}

bool ends_with(const std::string &s, const char *suffix)
{
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

}  // namespace

extern "C" int kf_load_calibration(const char *path, kf_cam_params *out)
{
    if (!path || !out) return kf_fail(KF_ERR_INVALID, "NULL argument");
    const std::string p(path);
    std::string doc;
    kf_cam_params c{};
    if (ends_with(p, ".xml")) {
        if (!read_file(path, doc)) return kf_fail(KF_ERR_INVALID, ("cannot read " + p).c_str());
        std::string cm, di;
        if (!xml_data(doc, "Camera_Matrix", cm) || !xml_data(doc, "Distortion_Coefficients", di))
            return kf_fail(KF_ERR_INVALID, ("no Camera_Matrix / Distortion_Coefficients data in " + p).c_str());
        const std::vector<float> m = split_numbers(cm), d = split_numbers(di);
        if (m.size() < 6 || d.size() < 5)
            return kf_fail(KF_ERR_INVALID, ("too few calibration values in " + p).c_str());
        c.fx = m[0];
        c.fy = m[4];
        c.ppx = m[2];
        c.ppy = m[5];
        c.k1 = d[0];
        c.k2 = d[1];
        c.p1 = d[2];
        c.p2 = d[3];
        c.k3 = d[4];
    } else if (ends_with(p, ".json")) {
        if (!read_file(path, doc)) return kf_fail(KF_ERR_INVALID, ("cannot read " + p).c_str());
        std::vector<double> m, d;
        if (!json_numbers(doc, "camera_matrix", m) || !json_numbers(doc, "dist_coeff", d) || m.size() < 6 ||
            d.size() < 5)
            return kf_fail(KF_ERR_INVALID, ("no camera_matrix / dist_coeff in " + p).c_str());
        c.fx = (float)m[0];  // [0][0]
        c.fy = (float)m[4];  // [1][1]
        c.ppx = (float)m[2]; // [0][2]
        c.ppy = (float)m[5]; // [1][2]
        c.k1 = (float)d[0];
        c.k2 = (float)d[1];
        c.p1 = (float)d[2];
        c.p2 = (float)d[3];
        c.k3 = (float)d[4];
    } else {
        return kf_fail(KF_ERR_INVALID, "please use a .xml or .json file for loading calibration data");
    }
    *out = c;
    return KF_OK;
}
