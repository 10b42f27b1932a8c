// Command-line driver with the reference's switches (main.cpp:56-79, :347-356):
//   lfrv_cli -config scene.cfg -algorithm IBR_gradient [-name value ...] [+flag]
// The configuration is parsed as coco::config does, written to <outdir>/config,
// and the algorithm switch selects IBR_gradient, fromMVE2coco (the warps of
// an MVE scene, computed on the GPU), or IBR_color, which is also what any
// other name runs (main.cpp:430-434).  The reference's interactive OpenGL
// algorithms (IBR_direct, IBR_laplacian, refocus, IBR_optical, IBR_video, and
// the scene viewer an empty name starts) are not part of this library and are
// reported as such.
#include <cstdio>
#include <exception>
#include <set>
#include <string>

#include "../../include/lfrv_ref.hpp"

int main(int argc, char **argv)
{
    using namespace lfrv_ref;
    Config cfg;
    cfg.parse_command_line(argc, argv);
    ConfigData cd;
    cd.parse_config(cfg);
    cfg.dump(cd._outdir + "/config");

    std::string algo;
    cfg.get_switch("algorithm", algo);
    try {
        static const std::set<std::string> opengl = {"IBR_direct", "IBR_laplacian", "refocus", "IBR_optical",
                                                     "IBR_video", ""};
        if (algo == "IBR_gradient") {
            IBR_gradient(&cd);
        } else if (algo == "fromMVE2coco") {
            from_mve_to_coco(&cd);
        } else if (opengl.count(algo)) {
            std::fprintf(stderr, "lfrv_cli: algorithm '%s' is an OpenGL path of the reference and is not provided "
                                 "(IBR_gradient, IBR_color, fromMVE2coco)\n", algo.c_str());
            return 2;
        } else {
            sr_synthesize_view(&cd);  // IBR_color, and the reference's default branch
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "lfrv_cli: %s\n", e.what());
        return 1;
    }
    return 0;
}
