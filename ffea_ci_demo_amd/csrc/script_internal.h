// Library-internal access to a loaded script's output and checkpoint file names, for the
// World-loop driver (ffea_world.cpp). Not part of the C ABI.
#pragma once
#include <string>

#include "../../include/ffea_script.h"

struct ScriptFiles {
    std::string script;          // the .ffea path as given
    std::string trajectory;      // trajectory_out_fname, or <dir>/<stem>/.ftj (SimulationParams.cpp:972-975)
    std::string measurement;     // measurement_out_fname, or <dir>/<stem>/.fm (:977-980)
    std::string checkpoint_in;   // checkpoint_in ("" if not set)
    std::string checkpoint_out;  // checkpoint_out, or the script path with .fcp (:984-990)
    std::string detailed;        // det_measurement_out_fname, or the measurement file with .fdm for its
                                 // extension when measurement_out_fname is given (:690-709); "" otherwise
};

bool ffea_internal_script_files(const ffea_script *s, ScriptFiles *out);
