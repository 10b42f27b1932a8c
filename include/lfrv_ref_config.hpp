// Configuration front end with the reference's semantics.
//
// Config mirrors coco::config (cocolib/common/parse_config.cpp:43-196): a
// map of switches read from "name value" lines of a config file or from
// "-name value" / "+flag" command-line arguments, where the switch "config"
// includes another file at that point.  ConfigData mirrors Config_data
// (sources/config.h:11-122, sources/config.cpp:11-128): the same field names,
// defaults and parsing, including its quirks (sigma_sensor is always divided
// by 255; outfile is prefixed by outdir).
#pragma once

#include <map>
#include <string>
#include <vector>

namespace lfrv_ref {

class Config {
public:
    bool parse_config_file(const std::string &filename);
    bool parse_command_line(int argc, char **argv);
    void set_switch(const std::string &name, const std::string &value);
    void add_string(const std::string &s) { strings_.push_back(s); }

    bool get_switch(const std::string &name) const;
    bool get_switch(const std::string &name, std::string &value) const;
    bool get_switch(const std::string &name, int &value) const;
    bool get_switch(const std::string &name, unsigned int &value) const;
    bool get_switch(const std::string &name, float &value) const;
    bool get_switch(const std::string &name, double &value) const;

    const std::vector<std::string> &strings() const { return strings_; }
    const std::map<std::string, std::string> &switches() const { return switches_; }
    bool dump(const std::string &filename) const;

private:
    std::map<std::string, std::string> switches_;
    std::vector<std::string> strings_;
};

struct ConfigData {
    int _unitTest = 0;
    size_t _nviews = 0;
    int _s_min = -1, _s_max = -1, _s_rmv = -1;
    int _t_min = 1, _t_max = 1, _t_rmv = -1;
    std::string _mve_name, _imageName, _cameraName, _flowAlg;
    std::string _lf_name, _init_name, _tau_name, _dpart_name, _uWarped_name, _uDepth_name;
    int _scale_min = -1, _scale_max = -1;
    std::string _depth_name, _depthFromPC_name, _depthFromMVE_name, _depthSuper_name;
    int _pyramidHeight = 0;
    std::string _cam_name, _ply_name, _patch_name;
    std::string _outdir, _outfile;
    int _dsf = 0;
    unsigned int _w = 0, _h = 0, _W = 0, _H = 0;
    int _windowW1 = 0, _windowW2 = 0, _windowH1 = 0, _windowH2 = 0;
    int _nchannels = 0;
    float _sigma_sensor = 1.0f / 255.0f;
    float _ugrad_threshold = 1.0f;
    float _lambda = 0.0f;
    float _lambda_max_factor = 0.0f;
    int _dw_type = 0, _gw_type = 0;
    float _dt_alpha = 1.0f, _dt_beta = 0.0f;
    int _niter = 0, _nmeta_iter = 1;
    float _gradient_step = 0.1f;
    // Not in the reference: which system possion_conj_grad solves.
    // "poisson" (default, the reference's): (lambda I + D^T D) u = lambda u0 + D^T g
    // with lambda = 1 (solver.cu:361-505, 541).  "normal_equations": the
    // gradient-constraint normal equations with the warped data term
    // (dt_alpha sum S^T W S + dt_beta D^T D), lfrv_cg_solve.
    std::string _ibr_system = "poisson";

    // Config_data::parse_config (config.cpp:67-128)
    void parse_config(const Config &cfg);
};

}  // namespace lfrv_ref
