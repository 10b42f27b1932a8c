// Configuration front end; see include/lfrv_ref_config.hpp.
#include "../../include/lfrv_ref_config.hpp"

#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <locale>
#include <sstream>

namespace lfrv_ref {

// One "name value..." line per switch; the value is the rest of the line from
// the first occurrence of its first word (parse_config.cpp:43-100).  Lines
// whose name starts with '#' are comments; a lone word is a global string.
bool Config::parse_config_file(const std::string &filename)
{
    FILE *f = std::fopen(filename.c_str(), "r");
    if (!f) {
        std::fprintf(stderr, "could not read config file %s\n", filename.c_str());
        return false;
    }
    char line[600];
    while (std::fgets(line, 500, f)) {
        size_t L = std::strlen(line);
        if (L > 0 && line[L - 1] == '\n') line[L - 1] = 0;
        char name[600] = {0}, value[600] = {0};
        const int count = std::sscanf(line, "%599s %599s", name, value);
        if (count < 1) continue;
        if (name[0] == '#') continue;
        std::string svalue;
        if (count > 1) {
            const char *loc = std::strstr(line, value);
            if (!loc) continue;
            svalue = loc;
        }
        if (!svalue.empty()) set_switch(name, svalue);
        else add_string(name);
    }
    std::fclose(f);
    return true;
}

// "-name value" sets a switch, "+name" a flag (value "1"), anything else is a
// global string (parse_config.cpp:103-157).
bool Config::parse_command_line(int argc, char **argv)
{
    int pos = 1;
    while (pos < argc) {
        const char *a = argv[pos];
        if (std::strlen(a) > 0) {
            if (a[0] == '-') {
                std::string v = "1";
                if (pos + 1 < argc) v = argv[pos + 1];
                set_switch(a + 1, v);
                pos++;
            } else if (a[0] == '+') {
                set_switch(a + 1, "1");
            } else {
                add_string(a);
            }
        }
        pos++;
    }
    return true;
}

// "config" includes a file where it appears (parse_config.cpp:178-192).
void Config::set_switch(const std::string &name, const std::string &value)
{
    if (name == "config") parse_config_file(value);
    switches_[name] = value;
}

bool Config::get_switch(const std::string &name) const { return switches_.count(name) != 0; }

bool Config::get_switch(const std::string &name, std::string &value) const
{
    auto it = switches_.find(name);
    if (it == switches_.end()) return false;
    value = it->second;
    return true;
}

bool Config::get_switch(const std::string &name, int &value) const
{
    auto it = switches_.find(name);
    if (it == switches_.end()) return false;
    value = std::atoi(it->second.c_str());
    return true;
}

template <class T>
static bool parse_c_locale(const std::string &s, T &value)
{
    std::istringstream is(s);
    is.imbue(std::locale("C"));
    is >> value;
    return true;
}

bool Config::get_switch(const std::string &name, unsigned int &value) const
{
    auto it = switches_.find(name);
    return it != switches_.end() && parse_c_locale(it->second, value);
}

bool Config::get_switch(const std::string &name, float &value) const
{
    auto it = switches_.find(name);
    return it != switches_.end() && parse_c_locale(it->second, value);
}

bool Config::get_switch(const std::string &name, double &value) const
{
    auto it = switches_.find(name);
    return it != switches_.end() && parse_c_locale(it->second, value);
}

bool Config::dump(const std::string &filename) const
{
    std::ofstream out(filename.c_str());
    if (!out) return false;
    for (const auto &kv : switches_) out << kv.first << " " << kv.second << "\n";
    return true;
}

void ConfigData::parse_config(const Config &cfg)
{
    cfg.get_switch("outdir", _outdir);
    if (!_outdir.empty()) ::mkdir(_outdir.c_str(), 0755);
    const size_t i = _outdir.rfind('/');
    if (i == std::string::npos || i != _outdir.length() - 1) _outdir += "/";
    cfg.get_switch("outfile", _outfile);
    _outfile = _outdir + _outfile;

    cfg.get_switch("unitTest", _unitTest);
    cfg.get_switch("s_min", _s_min);
    cfg.get_switch("s_max", _s_max);
    cfg.get_switch("s_rmv", _s_rmv);
    cfg.get_switch("t_min", _t_min);
    cfg.get_switch("t_max", _t_max);
    cfg.get_switch("t_rmv", _t_rmv);
    cfg.get_switch("mve_name", _mve_name);
    cfg.get_switch("imageName", _imageName);
    cfg.get_switch("cameraName", _cameraName);
    cfg.get_switch("flowAlg", _flowAlg);
    cfg.get_switch("lf_name", _lf_name);
    cfg.get_switch("init_name", _init_name);
    cfg.get_switch("tau_name", _tau_name);
    cfg.get_switch("dpart_name", _dpart_name);
    cfg.get_switch("uWarped_name", _uWarped_name);
    cfg.get_switch("uDepth_name", _uDepth_name);
    cfg.get_switch("scale_min", _scale_min);
    cfg.get_switch("scale_max", _scale_max);
    cfg.get_switch("depth_name", _depth_name);
    cfg.get_switch("depthFromPC_name", _depthFromPC_name);
    cfg.get_switch("depthFromMVE_name", _depthFromMVE_name);
    cfg.get_switch("depthSuper_name", _depthSuper_name);
    cfg.get_switch("pyramidHeight", _pyramidHeight);
    cfg.get_switch("cam_name", _cam_name);
    cfg.get_switch("ply_name", _ply_name);
    cfg.get_switch("patch_name", _patch_name);
    // outdir / outfile are read a second time as in config.cpp:104-105
    cfg.get_switch("outdir", _outdir);
    cfg.get_switch("outfile", _outfile);
    cfg.get_switch("dsf", _dsf);
    cfg.get_switch("width", _w);
    cfg.get_switch("height", _h);
    cfg.get_switch("window_width_1", _windowW1);
    cfg.get_switch("window_width_2", _windowW2);
    cfg.get_switch("window_height_1", _windowH1);
    cfg.get_switch("window_height_2", _windowH2);
    cfg.get_switch("nchannels", _nchannels);
    cfg.get_switch("ugrad_threshold", _ugrad_threshold);
    cfg.get_switch("sigma_sensor", _sigma_sensor);
    _sigma_sensor /= 255.0f;  // unconditionally, as config.cpp:118-119
    cfg.get_switch("lambda", _lambda);
    cfg.get_switch("lambda_max_factor", _lambda_max_factor);
    cfg.get_switch("dw_type", _dw_type);
    cfg.get_switch("gw_type", _gw_type);
    cfg.get_switch("dt_alpha", _dt_alpha);
    cfg.get_switch("dt_beta", _dt_beta);
    cfg.get_switch("niter", _niter);
    cfg.get_switch("nmeta_iter", _nmeta_iter);
    cfg.get_switch("gradient_step", _gradient_step);
    cfg.get_switch("ibr_system", _ibr_system);  // extension switch (not read by the reference)
}

}  // namespace lfrv_ref
