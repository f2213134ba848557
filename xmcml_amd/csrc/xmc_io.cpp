// xmcml's file formats (XML input, binary .surface, .mcml.out, .mcml.bk) and
// the backup/resume loop of its launcher, on the host.  The byte layouts are
// the reference's (base/xmcml/xmcmlLauncher/writer.cpp, reader.cpp,
// parser.cpp); the XML reader is this library's own small parser, which keeps
// the parts of tinyxml's behaviour the input files depend on: an element's
// text is the first child node when that node is text, surrounding white
// space is dropped, the five predefined entities and character references
// are decoded, and comments, declarations and CDATA are understood.
#include "../../include/xmcml_io.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace xmc {
int set_error(int code, const std::string& msg);
}

namespace {

using xmc::set_error;

// --------------------------------------------------------------------------
// XML
// --------------------------------------------------------------------------
struct XElem {
    std::string name;
    bool has_text = false;  // the first child node is text
    std::string text;
    std::vector<XElem> kids;
};

class XmlReader {
  public:
    explicit XmlReader(const std::string& s) : s_(s) {}

    bool parse_document(XElem& root, std::string& err) {
        if (s_.compare(0, 3, "\xEF\xBB\xBF") == 0) i_ = 3;
        // prolog: declarations, comments, doctype, white space
        while (true) {
            skip_ws();
            if (starts("<?")) {
                if (!skip_past("?>")) return fail("unterminated declaration", err);
            } else if (starts("<!--")) {
                if (!skip_past("-->")) return fail("unterminated comment", err);
            } else if (starts("<!")) {
                if (!skip_past(">")) return fail("unterminated doctype", err);
            } else {
                break;
            }
        }
        if (!starts("<")) return fail("no root element", err);
        return element(root, err, 0);
    }

  private:
    const std::string& s_;
    size_t i_ = 0;

    bool fail(const std::string& m, std::string& err) {
        err = "XML line " + std::to_string(line_at(i_)) + ": " + m;
        return false;
    }
    int line_at(size_t pos) const {
        int l = 1;
        for (size_t k = 0; k < pos && k < s_.size(); k++) l += s_[k] == '\n';
        return l;
    }
    bool starts(const char* p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
    bool skip_past(const char* p) {
        size_t k = s_.find(p, i_);
        if (k == std::string::npos) return false;
        i_ = k + std::strlen(p);
        return true;
    }
    static bool ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }
    void skip_ws() {
        while (i_ < s_.size() && ws(s_[i_])) i_++;
    }
    static bool name_char(char c) {
        return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_' || c == '-' ||
               c == '.' || c == ':' || (unsigned char)c >= 0x80;
    }
    std::string name() {
        size_t b = i_;
        while (i_ < s_.size() && name_char(s_[i_])) i_++;
        return s_.substr(b, i_ - b);
    }
    // entity and character references of a text run
    static bool decode(const std::string& in, std::string& out) {
        out.clear();
        for (size_t k = 0; k < in.size(); k++) {
            if (in[k] != '&') {
                out += in[k];
                continue;
            }
            size_t e = in.find(';', k);
            if (e == std::string::npos) return false;
            std::string ent = in.substr(k + 1, e - k - 1);
            if (ent == "amp") out += '&';
            else if (ent == "lt") out += '<';
            else if (ent == "gt") out += '>';
            else if (ent == "quot") out += '"';
            else if (ent == "apos") out += '\'';
            else if (ent.size() > 1 && ent[0] == '#') {
                unsigned long cp = (ent[1] == 'x' || ent[1] == 'X') ? std::strtoul(ent.c_str() + 2, nullptr, 16)
                                                                     : std::strtoul(ent.c_str() + 1, nullptr, 10);
                if (cp < 0x80) {
                    out += (char)cp;
                } else if (cp < 0x800) {
                    out += (char)(0xC0 | (cp >> 6));
                    out += (char)(0x80 | (cp & 0x3F));
                } else {
                    out += (char)(0xE0 | (cp >> 12));
                    out += (char)(0x80 | ((cp >> 6) & 0x3F));
                    out += (char)(0x80 | (cp & 0x3F));
                }
            } else {
                return false;
            }
            k = e;
        }
        return true;
    }
    static std::string condense(const std::string& t) {
        std::string r;
        bool gap = false;
        for (char c : t) {
            if (ws(c)) {
                gap = !r.empty();
            } else {
                if (gap) r += ' ';
                gap = false;
                r += c;
            }
        }
        return r;
    }

    bool element(XElem& e, std::string& err, int depth) {
        if (depth > 64) return fail("elements nested too deeply", err);
        i_++;  // '<'
        e.name = name();
        if (e.name.empty()) return fail("element without a name", err);
        // attributes are read and ignored (the input format has none)
        while (true) {
            skip_ws();
            if (i_ >= s_.size()) return fail("unterminated tag <" + e.name + ">", err);
            if (starts("/>")) {
                i_ += 2;
                return true;
            }
            if (s_[i_] == '>') {
                i_++;
                break;
            }
            std::string a = name();
            if (a.empty()) return fail("bad attribute in <" + e.name + ">", err);
            skip_ws();
            if (i_ >= s_.size() || s_[i_] != '=') return fail("attribute without value in <" + e.name + ">", err);
            i_++;
            skip_ws();
            if (i_ >= s_.size() || (s_[i_] != '"' && s_[i_] != '\'')) return fail("unquoted attribute", err);
            char q = s_[i_++];
            size_t k = s_.find(q, i_);
            if (k == std::string::npos) return fail("unterminated attribute", err);
            i_ = k + 1;
        }
        bool first = true;
        while (true) {
            if (i_ >= s_.size()) return fail("missing </" + e.name + ">", err);
            if (starts("</")) {
                i_ += 2;
                std::string n = name();
                skip_ws();
                if (n != e.name || i_ >= s_.size() || s_[i_] != '>')
                    return fail("</" + n + "> closes <" + e.name + ">", err);
                i_++;
                return true;
            }
            if (starts("<!--")) {
                if (!skip_past("-->")) return fail("unterminated comment", err);
                first = false;
                continue;
            }
            if (starts("<![CDATA[")) {
                i_ += 9;
                size_t k = s_.find("]]>", i_);
                if (k == std::string::npos) return fail("unterminated CDATA", err);
                if (first) {
                    e.has_text = true;
                    e.text = s_.substr(i_, k - i_);
                }
                i_ = k + 3;
                first = false;
                continue;
            }
            if (starts("<?")) {
                if (!skip_past("?>")) return fail("unterminated declaration", err);
                first = false;
                continue;
            }
            if (s_[i_] == '<') {
                e.kids.emplace_back();
                if (!element(e.kids.back(), err, depth + 1)) return false;
                first = false;
                continue;
            }
            size_t k = s_.find('<', i_);
            if (k == std::string::npos) return fail("missing </" + e.name + ">", err);
            std::string raw = s_.substr(i_, k - i_);
            i_ = k;
            std::string c = condense(raw);
            if (c.empty()) continue;  // white space between elements is not a node
            std::string d;
            if (!decode(c, d)) return fail("bad entity in <" + e.name + ">", err);
            if (first) {
                e.has_text = true;
                e.text = d;
            }
            first = false;
        }
    }
};

// --------------------------------------------------------------------------
// binary helpers
// --------------------------------------------------------------------------
struct File {
    FILE* f = nullptr;
    File(const char* name, const char* mode) { f = std::fopen(name, mode); }
    ~File() {
        if (f) std::fclose(f);
    }
    template <class T>
    bool put(const T* p, size_t n) {
        return n == 0 || std::fwrite(p, sizeof(T), n, f) == n;
    }
    template <class T>
    bool put1(const T& v) {
        return put(&v, 1);
    }
    template <class T>
    bool get(T* p, size_t n) {
        return n == 0 || std::fread(p, sizeof(T), n, f) == n;
    }
    template <class T>
    bool get1(T& v) {
        return get(&v, 1);
    }
    // Bytes between the read position and the end of the file (-1 if unknown).
    long long remaining() {
        long here = std::ftell(f);
        if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) return -1;
        long end = std::ftell(f);
        if (std::fseek(f, here, SEEK_SET) != 0 || end < here) return -1;
        return (long long)(end - here);
    }
};

bool put_zeros_f64(File& F, size_t n) {
    static const double z[512] = {};
    while (n > 0) {
        size_t k = n < 512 ? n : 512;
        if (!F.put(z, k)) return false;
        n -= k;
    }
    return true;
}

bool put_zeros_u64(File& F, size_t n) {
    static const uint64_t z[512] = {};
    while (n > 0) {
        size_t k = n < 512 ? n : 512;
        if (!F.put(z, k)) return false;
        n -= k;
    }
    return true;
}

bool header(File& F, uint32_t id, uint32_t len) { return F.put1(id) && F.put1(len); }

// The sections of WriteOutputToFile (writer.cpp:431-471); section lengths are
// 32-bit, as in the reference.
bool write_result_sections(File& F, const xmc_input* in, const xmc_output* out) {
    const int nd = out->numberOfDetectors;
    // 0x01 number of photons (writer.cpp:32-51)
    if (!header(F, XMC_SECTION_NUMBER_OF_PHOTONS, 8) || !F.put1(in->numberOfPhotons)) return false;
    // 0x02 area (:53-89)
    {
        const xmc_area* a = in->area;
        double d[6] = {a->corner.x, a->corner.y, a->corner.z, a->length.x, a->length.y, a->length.z};
        int p[3] = {a->partitionNumber.x, a->partitionNumber.y, a->partitionNumber.z};
        if (!header(F, XMC_SECTION_AREA, 6 * 8 + 3 * 4) || !F.put(d, 6) || !F.put(p, 3)) return false;
    }
    // 0x03 cube detectors: center, length (:91-125)
    {
        int n = in->numberOfCubeDetectors;
        if (!header(F, XMC_SECTION_CUBE_DETECTORS, (uint32_t)(4 + 48 * n)) || !F.put1(n)) return false;
        for (int i = 0; i < n; i++) {
            const xmc_cube_detector& c = in->cubeDetector[i];
            double b[6] = {c.center.x, c.center.y, c.center.z, c.length.x, c.length.y, c.length.z};
            if (!F.put(b, 6)) return false;
        }
    }
    // 0x0A ring detectors: center, radii (:127-160)
    {
        int n = in->numberOfRingDetectors;
        if (!header(F, XMC_SECTION_RING_DETECTORS, (uint32_t)(4 + 40 * n)) || !F.put1(n)) return false;
        for (int i = 0; i < n; i++) {
            const xmc_ring_detector& r = in->ringDetector[i];
            double b[5] = {r.center.x, r.center.y, r.center.z, r.smallRadius, r.bigRadius};
            if (!F.put(b, 5)) return false;
        }
    }
    // 0x04 specular reflectance (:162-181)
    if (!header(F, XMC_SECTION_SPECULAR_REFLECTANCE, 8) || !F.put1(out->specularReflectance)) return false;
    // 0x05 absorption, 0x0C scattering (:183-227)
    const int G = out->gridSize;
    for (int s = 0; s < 2; s++) {
        uint32_t id = s == 0 ? XMC_SECTION_ABSORPTION : XMC_SECTION_SCATTERING;
        const double* g = s == 0 ? out->absorption : out->scattering;
        if (!header(F, id, (uint32_t)(4 + (size_t)G * 8)) || !F.put1(G)) return false;
        if (g ? !F.put(g, (size_t)G) : !put_zeros_f64(F, (size_t)G)) return false;
    }
    // 0x06 detector weights (:229-252)
    {
        if (!header(F, XMC_SECTION_DETECTOR_WEIGHTS, (uint32_t)(4 + 8 * nd)) || !F.put1(nd)) return false;
        for (int i = 0; i < nd; i++)
            if (!F.put1(out->detectorInfo[i].weight)) return false;
    }
    // 0x07 detector trajectories: photons, size, voxel counts (:254-296)
    {
        uint32_t len = 4;
        for (int i = 0; i < nd; i++) len += 8 + 4 + (uint32_t)out->detectorInfo[i].trajectorySize * 8u;
        if (!header(F, XMC_SECTION_DETECTOR_TRAJECTORIES, len) || !F.put1(nd)) return false;
        for (int i = 0; i < nd; i++) {
            const xmc_detector_info& d = out->detectorInfo[i];
            if (!F.put1(d.numberOfPhotons) || !F.put1(d.trajectorySize)) return false;
            if (d.trajectory ? !F.put(d.trajectory, (size_t)d.trajectorySize)
                             : !put_zeros_u64(F, (size_t)d.trajectorySize))
                return false;
        }
    }
    // 0x08 time scale: start, finish, photons, weight, target weight (:298-356)
    {
        uint32_t len = 4;
        for (int i = 0; i < nd; i++) len += 4 + (uint32_t)out->detectorInfo[i].timeScaleSize * (4 * 8 + 8);
        if (!header(F, XMC_SECTION_DETECTOR_TIME_SCALE, len) || !F.put1(nd)) return false;
        for (int i = 0; i < nd; i++) {
            const xmc_detector_info& d = out->detectorInfo[i];
            if (!F.put1(d.timeScaleSize)) return false;
            for (int j = 0; j < d.timeScaleSize; j++) {
                const xmc_time_info& t = d.timeScale[j];
                if (!F.put1(t.timeStart) || !F.put1(t.timeFinish) || !F.put1(t.numberOfPhotons) ||
                    !F.put1(t.weight) || !F.put1(t.targetWeight))
                    return false;
            }
        }
    }
    // 0x0B detector ranges: other weight, target weight (:393-429)
    {
        if (!header(F, XMC_SECTION_DETECTOR_RANGES, (uint32_t)(4 + 16 * nd)) || !F.put1(nd)) return false;
        for (int i = 0; i < nd; i++) {
            double other = out->detectorInfo[i].weight - out->detectorInfo[i].targetWeight;
            double target = out->detectorInfo[i].targetWeight;
            if (!F.put1(other) || !F.put1(target)) return false;
        }
    }
    return true;
}

int check_result_args(const xmc_input* in, const xmc_output* out, const char* fileName) {
    if (!in || !out || !fileName) return set_error(XMC_ERR_ARG, "input/output/file name is NULL");
    if (!in->area) return set_error(XMC_ERR_ARG, "input has no area");
    if (out->numberOfDetectors > 0 && !out->detectorInfo) return set_error(XMC_ERR_ARG, "output has no detectorInfo");
    if (in->numberOfCubeDetectors > 0 && !in->cubeDetector) return set_error(XMC_ERR_ARG, "cube detectors are NULL");
    if (in->numberOfRingDetectors > 0 && !in->ringDetector) return set_error(XMC_ERR_ARG, "ring detectors are NULL");
    return XMC_OK;
}

// Section header of a reading pass; false when the id differs.
bool expect(File& F, uint32_t id, uint32_t* len = nullptr) {
    uint32_t got = 0, l = 0;
    if (!F.get1(got) || !F.get1(l) || got != id) return false;
    if (len) *len = l;
    return true;
}

bool skip_section(File& F) {  // SkipCurrentSection, reader.cpp:7-25
    uint32_t id = 0, len = 0;
    if (!F.get1(id) || !F.get1(len)) return false;
    return std::fseek(F.f, (long)len, SEEK_CUR) == 0;
}

}  // namespace

// --------------------------------------------------------------------------
// the document
// --------------------------------------------------------------------------
struct xmc_document {
    xmc_input in{};
    xmc_area area{};
    bool has_area = false;
    std::vector<xmc_layer> layers;
    std::vector<std::vector<int>> surface_ids;
    std::vector<xmc_surface> surfaces;
    std::vector<std::vector<xmc_double3>> vertices;
    std::vector<std::vector<xmc_int3>> triangles;
    std::vector<xmc_cube_detector> cubes;
    std::vector<xmc_ring_detector> rings;
    xmc_biasing bias{};
    std::vector<std::vector<double>> table_elements;  // weight tables read from a table file
    std::vector<xmc_weight_table> tables;

    void relink() {
        in.area = has_area ? &area : nullptr;
        for (size_t i = 0; i < layers.size(); i++) layers[i].surfaceId = surface_ids[i].data();
        in.layerInfo = layers.empty() ? nullptr : layers.data();
        for (size_t i = 0; i < surfaces.size(); i++) {
            surfaces[i].vertices = vertices[i].data();
            surfaces[i].triangles = triangles[i].data();
        }
        in.surface = surfaces.empty() ? nullptr : surfaces.data();
        in.cubeDetector = cubes.empty() ? nullptr : cubes.data();
        in.ringDetector = rings.empty() ? nullptr : rings.data();
        for (size_t t = 0; t < tables.size(); t++) tables[t].elements = table_elements[t].data();
        in.weightTable = tables.empty() ? nullptr : tables.data();
        in.numberOfWeightTables = (int)tables.size();
    }
};

namespace {

// The number conversions of parser.cpp: atof / atoi / atoll on the text.
struct Num {
    std::string err;
    const char* text(const XElem& e) {
        if (!e.has_text) {
            if (err.empty()) err = "<" + e.name + "> has no text";
            return "0";
        }
        return e.text.c_str();
    }
    double f(const XElem& e) { return std::atof(text(e)); }
    int i(const XElem& e) { return std::atoi(text(e)); }
    long long ll(const XElem& e) { return std::atoll(text(e)); }
};

void read_xyz(const XElem& e, Num& num, xmc_double3& v, bool with_atoi = false) {
    for (const XElem& c : e.kids) {
        double x = with_atoi ? (double)num.i(c) : 0.0;
        if (c.name == "X") v.x = with_atoi ? x : num.f(c);
        else if (c.name == "Y") v.y = with_atoi ? x : num.f(c);
        else if (c.name == "Z") v.z = with_atoi ? x : num.f(c);
    }
}

void read_ixyz(const XElem& e, Num& num, xmc_int3& v) {
    for (const XElem& c : e.kids) {
        if (c.name == "X") v.x = num.i(c);
        else if (c.name == "Y") v.y = num.i(c);
        else if (c.name == "Z") v.z = num.i(c);
    }
}

// ParseInputFile's element walk (parser.cpp:57-486).
int fill_document(const XElem& root, xmc_document& d) {
    Num num;
    xmc_input& in = d.in;
    std::memset(&in, 0, sizeof(in));
    // defaults, parser.cpp:20-50
    in.timeScaleSize = 1;
    in.startDirection.z = 1.0;
    in.terminationMode = XMC_TERMINATE_KILL;
    d.bias.weightIntegralPrecision = 128;
    d.bias.weightTablePrecision = 262144;
    d.bias.attractiveFactor = 1.0;
    bool have_layer_count = false, have_cube_count = false, have_ring_count = false;
    // detectors keyed by their number of filter layers; equal keys keep file order
    std::multimap<int, xmc_cube_detector> cube_map;
    std::multimap<int, xmc_ring_detector> ring_map;
    bool cube_section = false, ring_section = false;
    int declared_cubes = 0, declared_rings = 0;

    for (const XElem& node : root.kids) {
        const std::string& n = node.name;
        if (n == "NumberOfPhotons") {
            in.numberOfPhotons = (uint64_t)num.ll(node);
        } else if (n == "MinWeight") {
            in.minWeight = num.f(node);
        } else if (n == "TimeStart") {
            in.timeStart = num.f(node);
        } else if (n == "TimeFinish") {
            in.timeFinish = num.f(node);
        } else if (n == "TimeScaleSize") {
            in.timeScaleSize = num.i(node);
        } else if (n == "TargetRange") {
            for (const XElem& c : node.kids)
                if (c.name == "Layer") {
                    int l = num.i(c);
                    if (l >= 0 && l < XMC_MAX_LAYERS) in.targetRangeLayers[l] = 1;
                }
        } else if (n == "StartPosition") {
            read_xyz(node, num, in.startPosition);
        } else if (n == "StartDirection") {
            read_xyz(node, num, in.startDirection);
        } else if (n == "Area") {
            d.has_area = true;
            std::memset(&d.area, 0, sizeof(d.area));
            for (const XElem& c : node.kids) {
                if (c.name == "Corner") read_xyz(c, num, d.area.corner);
                else if (c.name == "Length") read_xyz(c, num, d.area.length);
                else if (c.name == "PartitionNumber") read_ixyz(c, num, d.area.partitionNumber);
            }
        } else if (n == "Biasing") {
            d.bias.useBiasing = 1;
            for (const XElem& c : node.kids) {
                if (c.name == "WeightIntegral") {
                    for (const XElem& w : c.kids) {
                        if (w.name == "IntegralPrecision") d.bias.weightIntegralPrecision = num.i(w);
                        if (w.name == "TablePrecision") d.bias.weightTablePrecision = num.i(w);
                        if (w.name == "AttractiveFactor") d.bias.attractiveFactor = num.f(w);
                    }
                } else if (c.name == "TargetPoint") {
                    read_xyz(c, num, d.bias.targetPoint, true);  // atoi, parser.cpp:233-241
                }
            }
        } else if (n == "NumberOfLayers") {
            in.numberOfLayers = num.i(node);
            have_layer_count = true;
        } else if (n == "Layers") {
            if (!have_layer_count || in.numberOfLayers < 0 || in.numberOfLayers > XMC_MAX_LAYERS)
                return set_error(XMC_ERR_ARG, "<Layers> needs a preceding <NumberOfLayers> in 0..8");
            d.layers.assign(in.numberOfLayers, xmc_layer{});
            d.surface_ids.assign(in.numberOfLayers, {});
            int i = 0;
            for (const XElem& lc : node.kids) {
                if (i >= in.numberOfLayers) return set_error(XMC_ERR_ARG, "number of layers invalid");
                if (lc.name == "Layer") {
                    xmc_layer& L = d.layers[i];
                    bool have_nsurf = false;
                    for (const XElem& c : lc.kids) {
                        if (c.name == "RefractiveIndex") L.refractiveIndex = num.f(c);
                        else if (c.name == "AbsorptionCoefficient") L.absorptionCoefficient = num.f(c);
                        else if (c.name == "ScatteringCoefficient") L.scatteringCoefficient = num.f(c);
                        else if (c.name == "Anisotropy") L.anisotropy = num.f(c);
                        else if (c.name == "NumberOfSurfaces") {
                            L.numberOfSurfaces = num.i(c);
                            have_nsurf = true;
                        } else if (c.name == "SurfaceIds") {
                            if (!have_nsurf || L.numberOfSurfaces < 0)
                                return set_error(XMC_ERR_ARG, "<SurfaceIds> needs a preceding <NumberOfSurfaces>");
                            std::vector<int>& ids = d.surface_ids[i];
                            ids.assign(L.numberOfSurfaces, 0);
                            int j = 0;
                            for (const XElem& s : c.kids) {
                                if (j >= L.numberOfSurfaces)
                                    return set_error(XMC_ERR_ARG, "more <Id> than NumberOfSurfaces in layer " +
                                                                      std::to_string(i));
                                if (s.name == "Id") ids[j] = num.i(s);
                                j++;
                            }
                        }
                    }
                }
                i++;
            }
        } else if (n == "NumberOfCubeDetectors") {
            declared_cubes = num.i(node);
            have_cube_count = true;
        } else if (n == "CubeDetectors") {
            if (!have_cube_count || declared_cubes < 0)
                return set_error(XMC_ERR_ARG, "<CubeDetectors> needs a preceding <NumberOfCubeDetectors>");
            cube_section = true;
            // one record reused across elements, as parser.cpp:323 declares it outside the loop
            xmc_cube_detector cd{};
            int i = 0;
            for (const XElem& dc : node.kids) {
                if (i >= declared_cubes) return set_error(XMC_ERR_ARG, "number of cube detectors invalid");
                if (dc.name == "CubeDetector") {
                    int nfilter = XMC_MAX_LAYERS;
                    for (int j = 0; j < XMC_MAX_LAYERS; j++) cd.filterLayers[j] = 1;
                    cd.targetAngle = 90.0;
                    for (const XElem& c : dc.kids) {
                        if (c.name == "Center") read_xyz(c, num, cd.center);
                        else if (c.name == "Length") read_xyz(c, num, cd.length);
                        else if (c.name == "Filter") {
                            for (int j = 0; j < XMC_MAX_LAYERS; j++) cd.filterLayers[j] = 0;
                            nfilter = 0;
                            for (const XElem& f : c.kids)
                                if (f.name == "Layer") {
                                    int l = num.i(f);
                                    if (l < 0 || l >= XMC_MAX_LAYERS)
                                        return set_error(XMC_ERR_ARG, "FilterLayers invalid");
                                    cd.filterLayers[l] = 1;
                                    nfilter++;
                                }
                        } else if (c.name == "TargetAngle") {
                            cd.targetAngle = num.f(c);
                        }
                    }
                    cube_map.insert(std::make_pair(nfilter, cd));
                }
                i++;
            }
        } else if (n == "NumberOfRingDetectors") {
            declared_rings = num.i(node);
            have_ring_count = true;
        } else if (n == "RingDetectors") {
            if (!have_ring_count || declared_rings < 0)
                return set_error(XMC_ERR_ARG, "<RingDetectors> needs a preceding <NumberOfRingDetectors>");
            ring_section = true;
            xmc_ring_detector rd{};
            int i = 0;
            for (const XElem& dc : node.kids) {
                if (i >= declared_rings) return set_error(XMC_ERR_ARG, "number of ring detectors invalid");
                if (dc.name == "RingDetector") {
                    for (int j = 0; j < XMC_MAX_LAYERS; j++) rd.filterLayers[j] = 1;
                    int nfilter = XMC_MAX_LAYERS;
                    rd.targetAngle = 90.0;
                    for (const XElem& c : dc.kids) {
                        if (c.name == "Center") read_xyz(c, num, rd.center);
                        else if (c.name == "SmallRadius") rd.smallRadius = num.f(c);
                        else if (c.name == "BigRadius") rd.bigRadius = num.f(c);
                        else if (c.name == "Filter") {
                            for (int j = 0; j < XMC_MAX_LAYERS; j++) rd.filterLayers[j] = 0;
                            nfilter = 0;
                            int j = 0;
                            for (const XElem& f : c.kids) {
                                if (j >= XMC_MAX_LAYERS) return set_error(XMC_ERR_ARG, "number of FilterLayers invalid");
                                if (f.name == "Layer") {
                                    int l = num.i(f);
                                    if (l < 0 || l >= XMC_MAX_LAYERS)
                                        return set_error(XMC_ERR_ARG, "FilterLayers invalid");
                                    rd.filterLayers[l] = 1;
                                    nfilter++;
                                }
                                j++;
                            }
                        } else if (c.name == "TargetAngle") {
                            rd.targetAngle = num.f(c);
                        }
                    }
                    ring_map.insert(std::make_pair(nfilter, rd));
                }
                i++;
            }
        } else if (n == "NumberOfSurfaces") {
            in.numberOfSurfaces = num.i(node);
        }
        if (!num.err.empty()) return set_error(XMC_ERR_ARG, num.err);
    }
    if (!num.err.empty()) return set_error(XMC_ERR_ARG, num.err);
    if (!d.has_area) return set_error(XMC_ERR_ARG, "input has no <Area>");
    if (!have_layer_count || (int)d.layers.size() != in.numberOfLayers)
        return set_error(XMC_ERR_ARG, "input has no <Layers> for its <NumberOfLayers>");
    // parser.cpp:489-509: detectors in key order fill the declared count
    if (declared_cubes > 0 && (!cube_section || (int)cube_map.size() != declared_cubes))
        return set_error(XMC_ERR_ARG, "fewer <CubeDetector> elements than NumberOfCubeDetectors");
    if (declared_rings > 0 && (!ring_section || (int)ring_map.size() != declared_rings))
        return set_error(XMC_ERR_ARG, "fewer <RingDetector> elements than NumberOfRingDetectors");
    in.useBiasing = d.bias.useBiasing;
    in.attractiveFactor = d.bias.attractiveFactor;
    in.targetPoint = d.bias.targetPoint;
    in.weightIntegralPrecision = d.bias.weightIntegralPrecision;
    in.weightTablePrecision = d.bias.weightTablePrecision;
    for (auto& kv : cube_map) d.cubes.push_back(kv.second);
    for (auto& kv : ring_map) d.rings.push_back(kv.second);
    in.numberOfCubeDetectors = (int)d.cubes.size();
    in.numberOfRingDetectors = (int)d.rings.size();
    if (in.numberOfSurfaces < 0) return set_error(XMC_ERR_ARG, "NumberOfSurfaces < 0");
    d.relink();
    return XMC_OK;
}

bool read_all(const char* name, std::string& s) {
    File F(name, "rb");
    if (!F.f) return false;
    char buf[65536];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof(buf), F.f)) > 0) s.append(buf, k);
    return true;
}

}  // namespace

extern "C" {

int xmc_read_input_file(const char* fileName, xmc_document** doc) {
    if (!fileName || !doc) return set_error(XMC_ERR_ARG, "file name / document is NULL");
    *doc = nullptr;
    std::string text;
    if (!read_all(fileName, text)) return set_error(XMC_ERR_ARG, std::string("cannot read ") + fileName);
    XElem root;
    std::string err;
    XmlReader rd(text);
    if (!rd.parse_document(root, err)) return set_error(XMC_ERR_ARG, std::string(fileName) + ": " + err);
    std::unique_ptr<xmc_document> d(new xmc_document());
    int rc = fill_document(root, *d);
    if (rc) return rc;
    *doc = d.release();
    return XMC_OK;
}

int xmc_read_surface_file(xmc_document* doc, const char* fileName) {  // ParseSurfaceFile, parser.cpp:555-587
    if (!doc || !fileName) return set_error(XMC_ERR_ARG, "document / file name is NULL");
    File F(fileName, "rb");
    if (!F.f) return set_error(XMC_ERR_ARG, std::string("cannot open ") + fileName);
    uint32_t id = 0;
    int n = 0;
    if (!F.get1(id) || id != XMC_SECTION_SURFACES) return set_error(XMC_ERR_ARG, "not a surface file (section id)");
    if (!F.get1(n) || n != doc->in.numberOfSurfaces)
        return set_error(XMC_ERR_ARG, "surface count differs from the input's NumberOfSurfaces");
    std::vector<xmc_surface> surfaces(n);
    std::vector<std::vector<xmc_double3>> verts(n);
    std::vector<std::vector<xmc_int3>> tris(n);
    for (int i = 0; i < n; i++) {
        int nv = 0, nt = 0;
        if (!F.get1(nv) || nv < 0) return set_error(XMC_ERR_ARG, "surface " + std::to_string(i) + ": bad vertex count");
        verts[i].resize(nv);
        if (!F.get(verts[i].data(), (size_t)nv)) return set_error(XMC_ERR_ARG, "truncated surface file");
        if (!F.get1(nt) || nt < 0)
            return set_error(XMC_ERR_ARG, "surface " + std::to_string(i) + ": bad triangle count");
        tris[i].resize(nt);
        if (!F.get(tris[i].data(), (size_t)nt)) return set_error(XMC_ERR_ARG, "truncated surface file");
        surfaces[i].numberOfVertices = nv;
        surfaces[i].numberOfTriangles = nt;
    }
    doc->surfaces.swap(surfaces);
    doc->vertices.swap(verts);
    doc->triangles.swap(tris);
    doc->relink();
    return XMC_OK;
}

xmc_input* xmc_document_input(xmc_document* doc) { return doc ? &doc->in : nullptr; }

int xmc_document_biasing(const xmc_document* doc, xmc_biasing* b) {
    if (!doc || !b) return set_error(XMC_ERR_ARG, "document / biasing is NULL");
    *b = doc->bias;
    return XMC_OK;
}

void xmc_document_free(xmc_document* doc) { delete doc; }

int xmc_write_surface_file(const xmc_surface* s, int n, const char* fileName) {
    if ((!s && n > 0) || n < 0 || !fileName) return set_error(XMC_ERR_ARG, "surfaces / file name");
    File F(fileName, "wb");
    if (!F.f) return set_error(XMC_ERR_ARG, std::string("cannot create ") + fileName);
    uint32_t id = XMC_SECTION_SURFACES;
    bool ok = F.put1(id) && F.put1(n);
    for (int i = 0; ok && i < n; i++)
        ok = F.put1(s[i].numberOfVertices) && F.put(s[i].vertices, (size_t)s[i].numberOfVertices) &&
             F.put1(s[i].numberOfTriangles) && F.put(s[i].triangles, (size_t)s[i].numberOfTriangles);
    if (!ok || std::fflush(F.f) != 0) return set_error(XMC_ERR_ARG, std::string("write failed: ") + fileName);
    return XMC_OK;
}

int xmc_write_output_file(const xmc_input* in, const xmc_output* out, const char* fileName) {
    int rc = check_result_args(in, out, fileName);
    if (rc) return rc;
    File F(fileName, "wb");
    if (!F.f) return set_error(XMC_ERR_ARG, std::string("cannot create ") + fileName);
    if (!write_result_sections(F, in, out) || std::fflush(F.f) != 0)
        return set_error(XMC_ERR_ARG, std::string("write failed: ") + fileName);
    return XMC_OK;
}

int xmc_backup_file_name(uint64_t n, char* buf, int size) {  // NumberToString, writer.cpp:8-29
    if (!buf || size < 32) return set_error(XMC_ERR_ARG, "buffer too small");
    char num[24];
    if (n >= 1000000000ull) std::snprintf(num, sizeof(num), "%.2f%c", (double)((float)n / 1000000000.0f), 'B');
    else if (n >= 1000000ull) std::snprintf(num, sizeof(num), "%.2f%c", (double)((float)n / 1000000.0f), 'M');
    else if (n >= 1000ull) std::snprintf(num, sizeof(num), "%.2f%c", (double)((float)n / 1000.0f), 'T');
    else std::snprintf(num, sizeof(num), "%u", (unsigned)n);
    std::snprintf(buf, (size_t)size, "xmcml_%s.mcml.bk", num);
    return XMC_OK;
}

int xmc_write_backup_file(const xmc_input* in, const xmc_output* out, const xmc_rng* rng, const char* fileName) {
    char name[64];
    if (!rng) return set_error(XMC_ERR_ARG, "rng is NULL");
    if (!fileName) {
        if (!in) return set_error(XMC_ERR_ARG, "input is NULL");
        xmc_backup_file_name(in->numberOfPhotons, name, sizeof(name));
        fileName = name;
    }
    int rc = check_result_args(in, out, fileName);
    if (rc) return rc;
    File F(fileName, "wb");
    if (!F.f) return set_error(XMC_ERR_ARG, std::string("cannot create ") + fileName);
    // 0x09 random generator (writer.cpp:358-391): one (value, offset) record
    int threads = 1, procs = 1;
    bool ok = write_result_sections(F, in, out) &&
              header(F, XMC_SECTION_RANDOM_GENERATOR, (uint32_t)(2 * 4 + 2 * 8)) && F.put1(threads) &&
              F.put1(procs) && F.put1(rng->seed) && F.put1(rng->nextPhoton);
    if (!ok || std::fflush(F.f) != 0) return set_error(XMC_ERR_ARG, std::string("write failed: ") + fileName);
    return XMC_OK;
}

int xmc_read_backup_file(const char* fileName, xmc_output* out, xmc_rng* rng) {
    if (!fileName || !out) return set_error(XMC_ERR_ARG, "file name / output is NULL");
    if (out->numberOfDetectors > 0 && !out->detectorInfo) return set_error(XMC_ERR_ARG, "output has no detectorInfo");
    File F(fileName, "rb");
    if (!F.f) return set_error(XMC_ERR_ARG, std::string("cannot open ") + fileName);
    auto bad = [&](const char* what) { return set_error(XMC_ERR_ARG, std::string(fileName) + ": " + what); };
    const int G = out->gridSize, nd = out->numberOfDetectors;
    // ReadOutputFormBackupFile, reader.cpp:366-392; into temporaries first so
    // that a rejected file leaves the output untouched
    uint64_t photons = 0;
    if (!expect(F, XMC_SECTION_NUMBER_OF_PHOTONS) || !F.get1(photons)) return bad("number of photons section");
    for (int k = 0; k < 4; k++)  // area, cube detectors, ring detectors, specular
        if (!skip_section(F)) return bad("truncated header sections");
    std::vector<double> grids[2];
    for (int s = 0; s < 2; s++) {
        int g = 0;
        if (!expect(F, s == 0 ? XMC_SECTION_ABSORPTION : XMC_SECTION_SCATTERING) || !F.get1(g) || g != G)
            return bad(s == 0 ? "absorption section / grid size" : "scattering section / grid size");
        grids[s].resize((size_t)G);
        if (!F.get(grids[s].data(), (size_t)G)) return bad("truncated grid");
    }
    int n = 0;
    std::vector<double> weight(nd);
    if (!expect(F, XMC_SECTION_DETECTOR_WEIGHTS) || !F.get1(n) || n != nd || !F.get(weight.data(), (size_t)nd))
        return bad("detector weights section / detector count");
    std::vector<uint64_t> count(nd);
    std::vector<std::vector<uint64_t>> traj(nd);
    if (!expect(F, XMC_SECTION_DETECTOR_TRAJECTORIES) || !F.get1(n) || n != nd) return bad("trajectories section");
    for (int i = 0; i < nd; i++) {
        int ts = 0;
        if (!F.get1(count[i]) || !F.get1(ts) || ts != out->detectorInfo[i].trajectorySize)
            return bad("trajectory size");
        traj[i].resize((size_t)ts);
        if (!F.get(traj[i].data(), (size_t)ts)) return bad("truncated trajectories");
    }
    std::vector<std::vector<xmc_time_info>> bins(nd);
    if (!expect(F, XMC_SECTION_DETECTOR_TIME_SCALE) || !F.get1(n) || n != nd) return bad("time scale section");
    for (int i = 0; i < nd; i++) {
        const xmc_detector_info& d = out->detectorInfo[i];
        int T = 0;
        if (!F.get1(T) || T != d.timeScaleSize) return bad("time scale size");
        bins[i].resize((size_t)T);
        for (int j = 0; j < T; j++) {
            xmc_time_info& t = bins[i][j];
            if (!F.get1(t.timeStart) || !F.get1(t.timeFinish) || !F.get1(t.numberOfPhotons) || !F.get1(t.weight) ||
                !F.get1(t.targetWeight))
                return bad("truncated time scale");
            if (t.timeStart != d.timeScale[j].timeStart || t.timeFinish != d.timeScale[j].timeFinish)
                return bad("time bin edges differ from the output's");
        }
    }
    std::vector<double> target(nd);
    if (!expect(F, XMC_SECTION_DETECTOR_RANGES) || !F.get1(n) || n != nd) return bad("detector ranges section");
    for (int i = 0; i < nd; i++) {
        double other = 0.0;
        if (!F.get1(other) || !F.get1(target[i])) return bad("truncated detector ranges");
        if (weight[i] - (target[i] + other) > 1e-6) return bad("detector ranges disagree with the weights");
    }
    // ParseBackupFile, parser.cpp:589-611: the generator section
    xmc_rng r{0, 0};
    if (rng) {
        int threads = 0, procs = 0;
        if (!expect(F, XMC_SECTION_RANDOM_GENERATOR) || !F.get1(threads) || !F.get1(procs))
            return bad("random generator section");
        if (threads != 1 || procs != 1)
            return bad("random generator section holds per-thread MCG59 states, not (seed, next photon)");
        if (!F.get1(r.seed) || !F.get1(r.nextPhoton)) return bad("truncated random generator section");
    }
    out->numberOfPhotons = photons;
    for (int s = 0; s < 2; s++) {
        double* g = s == 0 ? out->absorption : out->scattering;
        if (g) std::memcpy(g, grids[s].data(), (size_t)G * sizeof(double));
    }
    for (int i = 0; i < nd; i++) {
        xmc_detector_info& d = out->detectorInfo[i];
        d.weight = weight[i];
        d.numberOfPhotons = count[i];
        if (d.trajectory) std::memcpy(d.trajectory, traj[i].data(), traj[i].size() * sizeof(uint64_t));
        for (int j = 0; j < d.timeScaleSize; j++) d.timeScale[j] = bins[i][j];
        d.targetWeight = target[i];
    }
    if (rng) *rng = r;
    return XMC_OK;
}

int xmc_write_weight_table_file(const xmc_input* in, const char* fileName) {
    // WriteWeightIntegralTableToFile, weight_integral.cpp:105-157
    if (!in || (in->numberOfWeightTables > 0 && !in->weightTable)) return set_error(XMC_ERR_ARG, "input / tables");
    if (!fileName) fileName = "xmcml_weight_integral_table.bin";
    File F(fileName, "wb");
    if (!F.f) return set_error(XMC_ERR_ARG, std::string("cannot create ") + fileName);
    uint32_t len = 8 + 4 + 4;  // GetWeightIntegralSectionLength, :159-171
    for (int t = 0; t < in->numberOfWeightTables; t++) len += 8 + 4 + (uint32_t)in->weightTable[t].numberOfElements * 8u;
    bool ok = header(F, XMC_SECTION_WEIGHT_INTEGRAL, len) && F.put1(in->attractiveFactor) &&
              F.put1(in->weightIntegralPrecision) && F.put1(in->numberOfWeightTables);
    for (int t = 0; ok && t < in->numberOfWeightTables; t++) {
        const xmc_weight_table& w = in->weightTable[t];
        ok = F.put1(w.anisotropy) && F.put1(w.numberOfElements) && F.put(w.elements, (size_t)w.numberOfElements);
    }
    if (!ok || std::fflush(F.f) != 0) return set_error(XMC_ERR_ARG, std::string("write failed: ") + fileName);
    return XMC_OK;
}

int xmc_read_weight_table_file(xmc_document* doc, const char* fileName) {
    // ParseWeightIntegralFile, weight_integral.cpp:173-247: the file must have been made for this
    // input's attractive factor, integral and table precisions, and hold every distinct anisotropy
    if (!doc || !fileName) return set_error(XMC_ERR_ARG, "document / file name is NULL");
    File F(fileName, "rb");
    if (!F.f) return set_error(XMC_ERR_ARG, std::string("cannot open ") + fileName);
    auto bad = [&](const std::string& what) { return set_error(XMC_ERR_ARG, std::string(fileName) + ": " + what); };
    const xmc_input& in = doc->in;
    if (!expect(F, XMC_SECTION_WEIGHT_INTEGRAL)) return bad("not a weight table file (section id)");
    double factor = 0.0;
    int precision = 0, n = 0;
    if (!F.get1(factor) || factor != in.attractiveFactor) return bad("attractive factor differs from the input's");
    if (!F.get1(precision) || precision != in.weightIntegralPrecision)
        return bad("integral precision differs from the input's");
    std::vector<double> gs;  // GetDifferentAnisotropies, :67-90
    for (int l = 1; l < in.numberOfLayers; l++) {
        double g = in.layerInfo[l].anisotropy;
        if (std::find(gs.begin(), gs.end(), g) == gs.end()) gs.push_back(g);
    }
    if (!F.get1(n) || n < (int)gs.size()) return bad("fewer tables than distinct anisotropies");
    if ((long long)n * 12 > F.remaining()) return bad("truncated table list");  // before sizing from a count
    std::vector<std::vector<double>> elems((size_t)n);
    std::vector<xmc_weight_table> tabs((size_t)n);
    for (int t = 0; t < n; t++) {
        int m = 0;
        if (!F.get1(tabs[t].anisotropy) || !F.get1(m)) return bad("truncated table header");
        if (m - 1 != in.weightTablePrecision) return bad("table size differs from the input's table precision");
        if (m < 1 || (long long)m * 8 > F.remaining()) return bad("truncated table");
        elems[t].resize((size_t)m);
        if (!F.get(elems[t].data(), (size_t)m)) return bad("truncated table");
        tabs[t].numberOfElements = m;
    }
    for (double g : gs) {
        bool found = false;
        for (const xmc_weight_table& w : tabs) found = found || std::fabs(w.anisotropy - g) < 1e-6;
        if (!found) return bad("no table for anisotropy " + std::to_string(g));
    }
    doc->table_elements.swap(elems);
    doc->tables.swap(tabs);
    doc->relink();
    return XMC_OK;
}

int xmc_launch_with_backups(const xmc_input* input, xmc_output* out, xmc_rng* rng, const xmc_options* opt,
                            uint64_t portion, const char* backupFileName) {
    if (!input || !out || !rng) return set_error(XMC_ERR_ARG, "input/output/rng is NULL");
    if (portion == 0) return set_error(XMC_ERR_ARG, "portionSize must be > 0");
    xmc_options o;
    std::memset(&o, 0, sizeof(o));
    if (opt) o = *opt;
    if (!out->absorption) o.skipAbsorption = 1;
    if (out->scattering) o.scattering = 1;
    for (int d = 0; d < out->numberOfDetectors; d++)
        if (out->detectorInfo && out->detectorInfo[d].trajectory) o.trajectories = 1;
    xmc_scene* sc = nullptr;
    int rc = xmc_scene_create(input, &o, &sc);
    if (rc) return rc;
    xmc_input in = *input;  // numberOfPhotons is rewritten per portion, as LaunchOMPWithBackups does
    const uint64_t total = input->numberOfPhotons;
    const uint64_t portions = total > 0 ? (total - 1) / portion : 0;
    for (uint64_t i = 0; i < portions && rc == XMC_OK; i++) {
        rc = xmc_scene_run(sc, rng->seed, rng->nextPhoton, portion, nullptr);
        if (rc == XMC_OK) rc = xmc_scene_collect(sc, out);
        if (rc != XMC_OK) break;
        rng->nextPhoton += portion;
        // DoBackup, launcher_mpi.cpp:311-333
        out->numberOfPhotons += portion;
        in.numberOfPhotons = out->numberOfPhotons;
        rc = xmc_write_backup_file(&in, out, rng, backupFileName);
    }
    if (rc == XMC_OK) {
        uint64_t rest = total - portions * portion;
        rc = xmc_scene_run(sc, rng->seed, rng->nextPhoton, rest, nullptr);
        if (rc == XMC_OK) rc = xmc_scene_collect(sc, out);
        if (rc == XMC_OK) {
            rng->nextPhoton += rest;
            out->numberOfPhotons += rest;
        }
    }
    int rc2 = xmc_scene_destroy(sc);
    return rc ? rc : rc2;
}

}  // extern "C"
