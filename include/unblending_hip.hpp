/*
 * unblending_hip.hpp -- header-only C++ layer over the C interface of
 * libunblending.so, mirroring the reference's legacy image-level C++ API and
 * its LayerInfos file format:
 *
 *   compute_color_unmixing(ColorImage, layer_infos, opaque, concurrency)
 *       unblending/include/unblending/unblending.hpp:21-24, unblending.cpp:477-487
 *   perform_matte_refinement(ColorImage, layers, layer_infos, opaque, smooth, concurrency)
 *       unblending.hpp:27-32, unblending.cpp:489-503
 *   composite_layers(layers, comp_ops, modes)
 *       unblending.hpp:34-36, unblending.cpp:638-664 (here on the device)
 *   import_layer_infos(path) / export_layer_infos(layer_infos, directory)
 *       unblending.hpp:52-56, unblending/src/unblending_io.cpp:69-170
 *   retrieve_name / retrieve_by_name
 *       unblending/include/unblending/blend_mode.hpp:26-67
 *
 * Differences from the reference, by design: a ColorImage holds interleaved
 * straight-alpha RGBA doubles (the reference keeps four planes); colour
 * models are Gaussian (the only kind the passes accept); failures throw
 * std::invalid_argument (bad arguments), std::bad_alloc (memory) or
 * std::runtime_error (device and other failures) with the library's message.
 * Image files (the reference's Qt loaders) are left to the caller; see the
 * command-line tool for PPM/PFM input.
 */
#ifndef UNBLENDING_HIP_HPP
#define UNBLENDING_HIP_HPP

#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "unblending.h"
#include "unblending_hip.h"

namespace unblending {

enum class BlendMode : int32_t {
    Normal = UNBLENDING_NORMAL,
    Multiply = UNBLENDING_MULTIPLY,
    Screen = UNBLENDING_SCREEN,
    Overlay = UNBLENDING_OVERLAY,
    Darken = UNBLENDING_DARKEN,
    Lighten = UNBLENDING_LIGHTEN,
    ColorDodge = UNBLENDING_COLOR_DODGE,
    ColorBurn = UNBLENDING_COLOR_BURN,
    HardLight = UNBLENDING_HARD_LIGHT,
    SoftLight = UNBLENDING_SOFT_LIGHT,
    Difference = UNBLENDING_DIFFERENCE,
    Exclusion = UNBLENDING_EXCLUSION,
    LinearDodge = UNBLENDING_LINEAR_DODGE
};

// blend_mode.hpp:26-47
inline std::string retrieve_name(BlendMode mode)
{
    static const char* const names[] = {"Normal",    "Multiply",  "Screen",     "Overlay",   "Darken",
                                        "Lighten",   "ColorDodge", "ColorBurn", "HardLight", "SoftLight",
                                        "Difference", "Exclusion", "LinearDodge"};
    const int i = static_cast<int>(mode);
    if (i < 0 || i > 12) throw std::invalid_argument("Unknown blend mode");
    return names[i];
}

// blend_mode.hpp:49-67 (an unknown name is an error here, not Normal)
inline BlendMode retrieve_by_name(const std::string& name)
{
    for (int i = 0; i <= 12; ++i)
        if (retrieve_name(static_cast<BlendMode>(i)) == name) return static_cast<BlendMode>(i);
    throw std::invalid_argument("Unknown blend mode name: " + name);
}

// comp_op.hpp:9-30
struct CompOp {
    int X = 1, Y = 1, Z = 1;
    static CompOp SourceOver() { return CompOp{1, 1, 1}; }
    static CompOp Plus() { return CompOp{2, 1, 1}; }
    bool is_source_over() const { return X == 1 && Y == 1 && Z == 1; }
    bool is_plus() const { return X == 2 && Y == 1 && Z == 1; }
};

using Mat3 = std::array<double, 9>;  // row-major
using Vec3 = std::array<double, 3>;

// The closed-form 3x3 inverse by cofactors (the inverse the reference takes
// of a covariance with Eigen, unblending_io.cpp:163, color_model.cpp:139-146).
inline Mat3 inverse3(const Mat3& m)
{
    const double c00 = m[4] * m[8] - m[5] * m[7];
    const double c01 = m[5] * m[6] - m[3] * m[8];
    const double c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    if (!(std::fabs(det) > 0.0)) throw std::invalid_argument("Singular covariance matrix");
    const double id = 1.0 / det;
    return Mat3{c00 * id, (m[2] * m[7] - m[1] * m[8]) * id, (m[1] * m[5] - m[2] * m[4]) * id,
                c01 * id, (m[0] * m[8] - m[2] * m[6]) * id, (m[2] * m[3] - m[0] * m[5]) * id,
                c02 * id, (m[1] * m[6] - m[0] * m[7]) * id, (m[0] * m[4] - m[1] * m[3]) * id};
}

// A Gaussian colour model: mean and precision (inverse covariance),
// color_model.hpp:19-59.
class GaussianColorModel {
public:
    GaussianColorModel() : mu_{0.5, 0.5, 0.5}, sigma_inv_{1, 0, 0, 0, 1, 0, 0, 0, 1} {}
    GaussianColorModel(const Vec3& mu, const Mat3& sigma_inv) : mu_(mu), sigma_inv_(sigma_inv) {}
    static GaussianColorModel from_sigma(const Vec3& mu, const Mat3& sigma)
    {
        return GaussianColorModel(mu, inverse3(sigma));
    }
    const Vec3& get_mu() const { return mu_; }
    const Mat3& get_sigma_inv() const { return sigma_inv_; }
    Mat3 get_sigma() const { return inverse3(sigma_inv_); }
    void set_sigma(const Mat3& sigma) { sigma_inv_ = inverse3(sigma); }

private:
    Vec3 mu_;
    Mat3 sigma_inv_;
};

// layer_info.hpp:10-40
struct LayerInfo {
    CompOp comp_op;
    BlendMode blend_mode = BlendMode::Normal;
    GaussianColorModel color_model;
};

inline std::vector<BlendMode> extract_blend_modes(const std::vector<LayerInfo>& infos)
{
    std::vector<BlendMode> out;
    for (const auto& i : infos) out.push_back(i.blend_mode);
    return out;
}

inline std::vector<CompOp> extract_comp_ops(const std::vector<LayerInfo>& infos)
{
    std::vector<CompOp> out;
    for (const auto& i : infos) out.push_back(i.comp_op);
    return out;
}

// An RGBA image of doubles, [row][column][RGBA] (straight alpha).
class ColorImage {
public:
    ColorImage() = default;
    ColorImage(int width, int height) : w_(width), h_(height), rgba_(4 * size_t(width) * size_t(height), 0.0)
    {
        if (width <= 0 || height <= 0) throw std::invalid_argument("Invalid image dimensions");
    }
    int width() const { return w_; }
    int height() const { return h_; }
    std::array<double, 4> get_rgba(int x, int y) const
    {
        const double* p = &rgba_[4 * (size_t(y) * w_ + x)];
        return {p[0], p[1], p[2], p[3]};
    }
    Vec3 get_rgb(int x, int y) const
    {
        const double* p = &rgba_[4 * (size_t(y) * w_ + x)];
        return {p[0], p[1], p[2]};
    }
    void set_rgba(int x, int y, const std::array<double, 4>& v)
    {
        double* p = &rgba_[4 * (size_t(y) * w_ + x)];
        for (int c = 0; c < 4; ++c) p[c] = v[c];
    }
    const double* data() const { return rgba_.data(); }
    double* data() { return rgba_.data(); }

private:
    int w_ = 0, h_ = 0;
    std::vector<double> rgba_;
};

namespace detail {

inline void throw_status(int32_t status)
{
    if (status == UNBLENDING_OK) return;
    const std::string msg = unblending_hip_last_error();
    switch (status) {
    case UNBLENDING_INVALID_ARGUMENT:
        throw std::invalid_argument(msg.empty() ? "invalid argument" : msg);
    case UNBLENDING_OUT_OF_MEMORY:
        throw std::bad_alloc();
    case UNBLENDING_DEVICE_ERROR:
        throw std::runtime_error(msg.empty() ? "device error" : msg);
    default:
        throw std::runtime_error(msg.empty() ? "unblending failed" : msg);
    }
}

inline std::vector<unblending_layer> c_layers(const std::vector<LayerInfo>& infos)
{
    std::vector<unblending_layer> out(infos.size());
    for (size_t i = 0; i < infos.size(); ++i) {
        const LayerInfo& l = infos[i];
        out[i].comp_x = l.comp_op.X;
        out[i].comp_y = l.comp_op.Y;
        out[i].comp_z = l.comp_op.Z;
        out[i].blend_mode = static_cast<int32_t>(l.blend_mode);
        for (int c = 0; c < 3; ++c) out[i].mean[c] = l.color_model.get_mu()[c];
        for (int k = 0; k < 9; ++k) out[i].precision[k] = l.color_model.get_sigma_inv()[k];
    }
    return out;
}

// RGB of an image as the packed view the C interface reads
inline std::vector<double> rgb_of(const ColorImage& image)
{
    const size_t n = size_t(image.width()) * size_t(image.height());
    std::vector<double> rgb(3 * n);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) rgb[3 * i + c] = image.data()[4 * i + c];
    return rgb;
}

inline std::vector<ColorImage> unpack(const std::vector<double>& packed, int w, int h, size_t layers)
{
    std::vector<ColorImage> out;
    const size_t n = size_t(w) * size_t(h);
    for (size_t l = 0; l < layers; ++l) {
        ColorImage im(w, h);
        std::copy(packed.begin() + 4 * l * n, packed.begin() + 4 * (l + 1) * n, im.data());
        out.push_back(std::move(im));
    }
    return out;
}

inline std::vector<double> pack(const std::vector<ColorImage>& layers, int w, int h)
{
    const size_t n = size_t(w) * size_t(h);
    std::vector<double> packed(4 * n * layers.size());
    for (size_t l = 0; l < layers.size(); ++l) {
        if (layers[l].width() != w || layers[l].height() != h)
            throw std::invalid_argument("Layer dimensions differ from the image's");
        std::copy(layers[l].data(), layers[l].data() + 4 * n, packed.begin() + 4 * l * n);
    }
    return packed;
}

inline unblending_options options(bool opaque, bool smooth, int concurrency)
{
    unblending_options o = UNBLENDING_OPTIONS_INIT;
    o.opaque_background = opaque ? 1 : 0;
    o.smooth_background = smooth ? 1 : 0;
    o.target_concurrency = concurrency;
    return o;
}

}  // namespace detail

// unblending.hpp:21-24: the first pass; layers bottom first.
inline std::vector<ColorImage> compute_color_unmixing(const ColorImage& image,
                                                      const std::vector<LayerInfo>& layer_infos,
                                                      bool has_opaque_background, int target_concurrency = 0)
{
    const std::vector<double> rgb = detail::rgb_of(image);
    const std::vector<unblending_layer> layers = detail::c_layers(layer_infos);
    const unblending_rgb_view view{rgb.data(), rgb.size(), image.width(), image.height(), 0};
    const unblending_options opts = detail::options(has_opaque_background, false, target_concurrency);
    std::vector<double> out(4 * size_t(image.width()) * size_t(image.height()) * layers.size());
    detail::throw_status(unblending_unmix(&view, layers.data(), layers.size(), &opts, out.data(), out.size()));
    return detail::unpack(out, image.width(), image.height(), layers.size());
}

// unblending.hpp:27-32: the refinement pass (guided filter with the
// reference's radius and epsilon, then the re-solve).
inline std::vector<ColorImage> perform_matte_refinement(const ColorImage& image,
                                                        const std::vector<ColorImage>& layers,
                                                        const std::vector<LayerInfo>& layer_infos,
                                                        bool has_opaque_background, bool force_smooth_background,
                                                        int target_concurrency = 0)
{
    const std::vector<double> rgb = detail::rgb_of(image);
    const std::vector<unblending_layer> cl = detail::c_layers(layer_infos);
    const std::vector<double> first = detail::pack(layers, image.width(), image.height());
    const unblending_rgb_view view{rgb.data(), rgb.size(), image.width(), image.height(), 0};
    const unblending_options opts =
        detail::options(has_opaque_background, force_smooth_background, target_concurrency);
    std::vector<double> out(first.size());
    detail::throw_status(unblending_refine(&view, cl.data(), cl.size(), first.data(), first.size(), &opts,
                                           out.data(), out.size()));
    return detail::unpack(out, image.width(), image.height(), cl.size());
}

// unblending.hpp:34-36: the layers composited bottom to top (on the device).
inline ColorImage composite_layers(const std::vector<ColorImage>& layers, const std::vector<CompOp>& comp_ops,
                                   const std::vector<BlendMode>& modes)
{
    if (layers.empty() || comp_ops.size() != layers.size() || modes.size() != layers.size())
        throw std::invalid_argument("One comp-op and one blend mode per layer are required");
    const int w = layers.front().width(), h = layers.front().height();
    const std::vector<double> packed = detail::pack(layers, w, h);
    std::vector<int32_t> ops, md;
    for (size_t l = 0; l < layers.size(); ++l) {
        ops.insert(ops.end(), {comp_ops[l].X, comp_ops[l].Y, comp_ops[l].Z});
        md.push_back(static_cast<int32_t>(modes[l]));
    }
    ColorImage out(w, h);
    detail::throw_status(unblending_hip_composite(packed.data(), packed.size(), w, h, layers.size(), ops.data(),
                                                  md.data(), out.data(), 4 * size_t(w) * size_t(h)));
    return out;
}

// ---- LayerInfos files (unblending_io.cpp:69-170, LayerInfos.md) ----

namespace detail {

// A JSON value of the LayerInfos format: numbers, strings, arrays, objects.
struct Json {
    enum Kind { Null, Number, String, Array, Object } kind = Null;
    double number = 0.0;
    std::string string;
    std::vector<Json> items;
    std::map<std::string, Json> fields;
    const Json& operator[](const std::string& k) const
    {
        static const Json null;
        auto it = fields.find(k);
        return it == fields.end() ? null : it->second;
    }
};

class JsonReader {
public:
    explicit JsonReader(const std::string& text) : s_(text) {}
    Json parse()
    {
        Json v = value();
        space();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    [[noreturn]] void fail(const char* what) const
    {
        throw std::invalid_argument(std::string("LayerInfo parse error: ") + what + " at offset " + std::to_string(i_));
    }
    void space()
    {
        while (i_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[i_]))) ++i_;
    }
    bool take(char c)
    {
        space();
        if (i_ < s_.size() && s_[i_] == c) {
            ++i_;
            return true;
        }
        return false;
    }
    std::string str()
    {
        if (!take('"')) fail("expected a string");
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\' && i_ < s_.size()) {
                const char e = s_[i_++];
                c = e == 'n' ? '\n' : e == 't' ? '\t' : e == 'r' ? '\r' : e == 'b' ? '\b' : e == 'f' ? '\f' : e;
            }
            out.push_back(c);
        }
        if (!take('"')) fail("unterminated string");
        return out;
    }
    Json value()
    {
        space();
        if (i_ >= s_.size()) fail("unexpected end");
        Json v;
        const char c = s_[i_];
        if (c == '{') {
            ++i_;
            v.kind = Json::Object;
            if (take('}')) return v;
            do {
                const std::string k = str();
                if (!take(':')) fail("expected ':'");
                v.fields[k] = value();
            } while (take(','));
            if (!take('}')) fail("expected '}'");
        } else if (c == '[') {
            ++i_;
            v.kind = Json::Array;
            if (take(']')) return v;
            do v.items.push_back(value());
            while (take(','));
            if (!take(']')) fail("expected ']'");
        } else if (c == '"') {
            v.kind = Json::String;
            v.string = str();
        } else if (s_.compare(i_, 4, "null") == 0) {
            i_ += 4;
        } else {
            size_t used = 0;
            try {
                v.number = std::stod(s_.substr(i_), &used);
            } catch (const std::exception&) {
                fail("expected a value");
            }
            v.kind = Json::Number;
            i_ += used;
        }
        return v;
    }
    const std::string& s_;
    size_t i_ = 0;
};

// a number as the reference's JSON writer prints it (%.17g)
inline std::string json_number(double v)
{
    if (!std::isfinite(v)) return "null";
    char buf[32];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

inline Vec3 json_vec3(const Json& j)
{
    if (j.kind != Json::Array || j.items.size() != 3) throw std::invalid_argument("LayerInfo parse error: primary_color");
    return Vec3{j.items[0].number, j.items[1].number, j.items[2].number};
}

// unblending_io.cpp:69-85: a number is a uniform variance, nine numbers a
// covariance stored column by column
inline Mat3 json_variance(const Json& j)
{
    if (j.kind == Json::Number) return Mat3{j.number, 0, 0, 0, j.number, 0, 0, 0, j.number};
    if (j.kind == Json::Array && j.items.size() == 9) {
        Mat3 m{};
        for (int col = 0; col < 3; ++col)
            for (int row = 0; row < 3; ++row) m[3 * row + col] = j.items[col * 3 + row].number;
        return m;
    }
    throw std::invalid_argument("LayerInfo parse error: color_variance");
}

// unblending_io.cpp:26-39: a multiple of the identity is written as one
// number, anything else as nine numbers (the writer's element order,
// unblending_io.cpp:14-24)
inline std::string variance_json(const Mat3& s)
{
    bool uniform = s[0] != 0.0;
    for (int r = 0; r < 3 && uniform; ++r)
        for (int c = 0; c < 3 && uniform; ++c)
            uniform = std::fabs(s[3 * r + c] / s[0] - (r == c ? 1.0 : 0.0)) <= 1e-12;
    if (uniform) return json_number(s[0]);
    std::string out = "[";
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row) {
            if (col || row) out += ", ";
            out += json_number(s[3 * col + row]);
        }
    return out + "]";
}

inline std::string json_string(const std::string& s) { return "\"" + s + "\""; }

}  // namespace detail

// unblending_io.cpp:143-170: colour models read as covariances and stored as
// their inverses; "source-over" or else plus.
inline std::vector<LayerInfo> parse_layer_infos(const std::string& text)
{
    const detail::Json json = detail::JsonReader(text).parse();
    if (json.kind != detail::Json::Array) throw std::invalid_argument("LayerInfo parse error: expected an array");
    std::vector<LayerInfo> infos;
    for (const detail::Json& j : json.items) {
        LayerInfo info;
        info.comp_op = j["comp_op"].string == "source-over" ? CompOp::SourceOver() : CompOp::Plus();
        info.blend_mode = retrieve_by_name(j["mode"].string);
        const detail::Json& model = j["color_model"];
        info.color_model =
            GaussianColorModel::from_sigma(detail::json_vec3(model["primary_color"]),
                                           detail::json_variance(model["color_variance"]));
        infos.push_back(info);
    }
    return infos;
}

inline std::vector<LayerInfo> import_layer_infos(const std::string& input_file_path)
{
    std::ifstream in(input_file_path);
    if (!in) throw std::invalid_argument("Cannot open " + input_file_path);
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    return parse_layer_infos(text);
}

// unblending_io.cpp:87-141: the writer's key order (sorted) and separators
inline std::string layer_infos_json(const std::vector<LayerInfo>& infos)
{
    std::string out = "[";
    for (size_t i = 0; i < infos.size(); ++i) {
        const LayerInfo& l = infos[i];
        const Vec3& mu = l.color_model.get_mu();
        const std::string op = l.comp_op.is_plus() ? "plus" : l.comp_op.is_source_over() ? "source-over" : "unknown";
        if (i) out += ", ";
        out += "{\"color_model\": {\"color_variance\": " + detail::variance_json(l.color_model.get_sigma()) +
               ", \"primary_color\": [" + detail::json_number(mu[0]) + ", " + detail::json_number(mu[1]) + ", " +
               detail::json_number(mu[2]) + "]}, \"comp_op\": " + detail::json_string(op) +
               ", \"mode\": " + detail::json_string(retrieve_name(l.blend_mode)) + "}";
    }
    return out + "]";
}

// unblending_io.cpp:133-139: writes <directory>/layer_infos.json
inline void export_layer_infos(const std::vector<LayerInfo>& layer_infos, const std::string& output_directory_path)
{
    std::ofstream out(output_directory_path + "/layer_infos.json");
    if (!out) throw std::runtime_error("Cannot write " + output_directory_path + "/layer_infos.json");
    out << layer_infos_json(layer_infos);
}

}  // namespace unblending

#endif  // UNBLENDING_HIP_HPP
