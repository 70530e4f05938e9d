// unblending_cli.cpp -- command-line front end of the two passes without Qt.
//
// The reference's unblending-cli (unblending-cli/main.cpp:13-96) for images
// in the netpbm family: reads an RGB image (PPM P6, 8 or 16 bits, or PFM
// "PF", 32-bit float) and a LayerInfos file, runs compute_color_unmixing and
// perform_matte_refinement with an opaque, smoothed background (main.cpp:59-66),
// and writes, into the output directory:
//   layer_<i>[_<Mode>].pam          refined layers, RGBA, 8 bits (PAM RGB_ALPHA)
//   layer-alpha_<i>.pgm             their alphas (with -v)
//   non-smoothed-layer_<i>[...].pam first-pass layers (with -v), with alphas
//   input.ppm, recomposited.ppm, non-smoothed-recomposited.ppm  (with -v)
//   layer_infos.json                the layer infos (unblending_io.cpp:133-139)
// Usage: unblending-cli [-o DIR] [-e] [-v] <input-image> <layer-infos.json>
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "unblending_hip.hpp"

using namespace unblending;

namespace {

int usage(const char* argv0, int code)
{
    std::cout << "unblending-cli - decompose an image into blended layers on the GPU\n"
              << "Usage: " << argv0 << " [-o DIR] [-e] [-v] <input-image-path> <layer-infos-path>\n"
              << "  -o, --outdir DIR               output directory (default ./out)\n"
              << "  -e, --explicit-mode-names      append blend mode names to layer file names\n"
              << "  -v, --verbose-export           also export the first pass, alphas and recomposites\n"
              << "  input image: PPM (P6, 8 or 16 bits) or PFM (PF, float)\n";
    return code;
}

// netpbm header token (skipping whitespace and # comments)
std::string token(std::istream& in)
{
    std::string t;
    int c;
    while ((c = in.get()) != EOF) {
        if (c == '#') {
            while ((c = in.get()) != EOF && c != '\n') {
            }
        } else if (!std::isspace(c)) {
            t.push_back(char(c));
            break;
        }
    }
    while ((c = in.peek()) != EOF && !std::isspace(c)) t.push_back(char(in.get()));
    return t;
}

ColorImage read_image(const std::string& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    const std::string magic = token(in);
    const int w = std::stoi(token(in)), h = std::stoi(token(in));
    ColorImage image(w, h);
    if (magic == "P6") {
        const int maxval = std::stoi(token(in));
        if (maxval <= 0 || maxval > 65535) throw std::runtime_error("unsupported PPM maxval");
        in.get();
        const int bytes = maxval < 256 ? 1 : 2;
        std::vector<unsigned char> raw(size_t(w) * h * 3 * bytes);
        if (!in.read(reinterpret_cast<char*>(raw.data()), std::streamsize(raw.size())))
            throw std::runtime_error("truncated PPM " + path);
        for (size_t i = 0; i < size_t(w) * h; ++i)
            for (int c = 0; c < 3; ++c) {
                const size_t k = 3 * i + c;
                const unsigned v = bytes == 1 ? raw[k] : (unsigned(raw[2 * k]) << 8) | raw[2 * k + 1];
                image.data()[4 * i + c] = double(v) / double(maxval);
            }
    } else if (magic == "PF") {
        const double scale = std::stod(token(in));
        in.get();
        std::vector<float> raw(size_t(w) * h * 3);
        if (!in.read(reinterpret_cast<char*>(raw.data()), std::streamsize(raw.size() * sizeof(float))))
            throw std::runtime_error("truncated PFM " + path);
        const bool little = scale < 0.0;
        const uint16_t probe = 1;
        const bool host_little = *reinterpret_cast<const unsigned char*>(&probe) == 1;
        for (size_t i = 0; i < raw.size(); ++i) {
            if (little != host_little) {
                unsigned char* b = reinterpret_cast<unsigned char*>(&raw[i]);
                std::swap(b[0], b[3]);
                std::swap(b[1], b[2]);
            }
        }
        for (int y = 0; y < h; ++y)  // PFM rows run bottom to top
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c)
                    image.data()[4 * (size_t(y) * w + x) + c] = raw[3 * (size_t(h - 1 - y) * w + x) + c];
    } else {
        throw std::runtime_error("unsupported image format (PPM P6 or PFM PF): " + path);
    }
    for (size_t i = 0; i < size_t(w) * h; ++i) image.data()[4 * i + 3] = 1.0;
    return image;
}

unsigned char byte_of(double v)
{
    const double c = std::min(1.0, std::max(0.0, v));
    return static_cast<unsigned char>(std::lround(c * 255.0));
}

void write_pam(const ColorImage& im, const std::string& path)
{
    std::ofstream out(path, std::ios::binary);
    out << "P7\nWIDTH " << im.width() << "\nHEIGHT " << im.height() << "\nDEPTH 4\nMAXVAL 255\nTUPLTYPE RGB_ALPHA\nENDHDR\n";
    const size_t n = size_t(im.width()) * im.height();
    std::vector<unsigned char> raw(4 * n);
    for (size_t i = 0; i < 4 * n; ++i) raw[i] = byte_of(im.data()[i]);
    out.write(reinterpret_cast<const char*>(raw.data()), std::streamsize(raw.size()));
}

void write_ppm(const ColorImage& im, const std::string& path)
{
    std::ofstream out(path, std::ios::binary);
    out << "P6\n" << im.width() << " " << im.height() << "\n255\n";
    const size_t n = size_t(im.width()) * im.height();
    std::vector<unsigned char> raw(3 * n);
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) raw[3 * i + c] = byte_of(im.data()[4 * i + c]);
    out.write(reinterpret_cast<const char*>(raw.data()), std::streamsize(raw.size()));
}

void write_alpha_pgm(const ColorImage& im, const std::string& path)
{
    std::ofstream out(path, std::ios::binary);
    out << "P5\n" << im.width() << " " << im.height() << "\n255\n";
    const size_t n = size_t(im.width()) * im.height();
    std::vector<unsigned char> raw(n);
    for (size_t i = 0; i < n; ++i) raw[i] = byte_of(im.data()[4 * i + 3]);
    out.write(reinterpret_cast<const char*>(raw.data()), std::streamsize(raw.size()));
}

// image_qt.cpp:65-83: <prefix>_<index>[_<Mode>] and <prefix>-alpha_<index>
void export_layers(const std::vector<ColorImage>& layers, const std::string& dir, const std::string& prefix,
                   bool with_alpha_channel, bool with_blend_mode_suffix, const std::vector<LayerInfo>& infos)
{
    for (size_t i = 0; i < layers.size(); ++i) {
        const std::string suffix = with_blend_mode_suffix ? "_" + retrieve_name(infos[i].blend_mode) : "";
        write_pam(layers[i], dir + "/" + prefix + "_" + std::to_string(i) + suffix + ".pam");
        if (with_alpha_channel) write_alpha_pgm(layers[i], dir + "/" + prefix + "-alpha_" + std::to_string(i) + ".pgm");
    }
}

void make_directory(const std::string& path)
{
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
                throw std::runtime_error("cannot create " + cur);
        }
        if (i < path.size()) cur.push_back(path[i]);
    }
}

}  // namespace

int main(int argc, char** argv)
{
    std::string outdir = "./out";
    bool explicit_names = false, verbose = false;
    std::vector<std::string> positional;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") return usage(argv[0], 0);
        if (a == "-e" || a == "--explicit-mode-names") explicit_names = true;
        else if (a == "-v" || a == "--verbose-export") verbose = true;
        else if ((a == "-o" || a == "--outdir") && i + 1 < argc) outdir = argv[++i];
        else if (a.size() > 1 && a[0] == '-') return usage(argv[0], 2);
        else positional.push_back(a);
    }
    if (positional.size() != 2) return usage(argv[0], 0);
    try {
        make_directory(outdir);
        const ColorImage image = read_image(positional[0]);
        const std::vector<LayerInfo> infos = import_layer_infos(positional[1]);

        constexpr bool has_opaque_background = true;    // main.cpp:59
        constexpr bool force_smooth_background = true;  // main.cpp:60
        const std::vector<ColorImage> layers = compute_color_unmixing(image, infos, has_opaque_background);
        const std::vector<ColorImage> refined =
            perform_matte_refinement(image, layers, infos, has_opaque_background, force_smooth_background);

        if (verbose) export_layers(layers, outdir, "non-smoothed-layer", true, explicit_names, infos);
        export_layers(refined, outdir, "layer", verbose, explicit_names, infos);
        if (verbose) {
            write_ppm(image, outdir + "/input.ppm");
            const auto modes = extract_blend_modes(infos);
            const auto ops = extract_comp_ops(infos);
            write_ppm(composite_layers(layers, ops, modes), outdir + "/non-smoothed-recomposited.ppm");
            write_ppm(composite_layers(refined, ops, modes), outdir + "/recomposited.ppm");
        }
        export_layer_infos(infos, outdir);
    } catch (const std::exception& e) {
        std::cerr << "unblending-cli: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
