// Test driver for the LayerInfos reader and writer of include/unblending_hip.hpp
// (no device work): `parse FILE` prints one line per layer
//   mode comp_x comp_y comp_z mu0 mu1 mu2 p00 .. p22 s00 .. s22
// (precision, then the covariance it implies, row-major, %.17g);
// `export FILE DIR` writes DIR/layer_infos.json from FILE's layers;
// `names` prints every blend mode's name and its id read back.
#include <cstdio>
#include <string>

#include "unblending_hip.hpp"

int main(int argc, char** argv)
{
    using namespace unblending;
    try {
        const std::string cmd = argc > 1 ? argv[1] : "";
        if (cmd == "parse" && argc == 3) {
            for (const LayerInfo& l : import_layer_infos(argv[2])) {
                std::printf("%d %d %d %d", int(l.blend_mode), l.comp_op.X, l.comp_op.Y, l.comp_op.Z);
                for (double v : l.color_model.get_mu()) std::printf(" %.17g", v);
                for (double v : l.color_model.get_sigma_inv()) std::printf(" %.17g", v);
                for (double v : l.color_model.get_sigma()) std::printf(" %.17g", v);
                std::printf("\n");
            }
            return 0;
        }
        if (cmd == "export" && argc == 4) {
            export_layer_infos(import_layer_infos(argv[2]), argv[3]);
            return 0;
        }
        if (cmd == "names") {
            for (int i = 0; i <= 12; ++i) {
                const std::string n = retrieve_name(static_cast<BlendMode>(i));
                std::printf("%s %d\n", n.c_str(), int(retrieve_by_name(n)));
            }
            return 0;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 2;
}
