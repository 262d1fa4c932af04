// TEST INFRASTRUCTURE ONLY: a driver that runs the Nimbus reference's own CM1
// input_sounding reader and file-sounding column builder (compiled from the
// reference sources where they lie, see the `ref` target of oracle/Makefile)
// to produce golden vectors for the sounding oracle and the backend's
// sounding entry points. It runs only in the development container; its
// outputs are committed as fixtures under tests/golden/.
//
// Usage:
//   ref_sounding read <file>
//       prints "level z p_hpa T theta qv u v" per level, or "error <message>"
//   ref_sounding column <file> <method> <extrap_below> <extrap_above> <dz> <z0> [<z1> ...]
//       prints "center z theta T p rho qv u v" per target height, or "error <message>"
// Numbers are printed with %.17g (exact round trip).
#include "init/sounding/file_sounding.hpp"
#include "soundings/schemes/cm1/cm1_sounding.hpp"

#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    const std::string op = argv[1], path = argv[2];
    try
    {
        if (op == "read")
        {
            SoundingConfig cfg;
            cfg.scheme_id = "cm1";
            cfg.file_path = path;
            cfg.use_fallback_profiles = false;
            Cm1SoundingScheme reader;
            reader.initialize(cfg);
            const SoundingData d = reader.load_sounding(path);
            for (std::size_t k = 0; k < d.num_levels(); ++k)
                std::printf("level %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", d.height_m[k], d.pressure_hpa[k],
                            d.temperature_k[k], d.potential_temperature_k[k], d.mixing_ratio_kgkg[k], d.wind_u_ms[k],
                            d.wind_v_ms[k]);
            return 0;
        }
        if (op == "column" && argc >= 8)
        {
            tmv::init::FileSoundingParams params;
            params.file_path = path;
            params.scheme_id = "cm1";
            params.interpolation_method = std::atof(argv[3]);
            params.extrapolate_below_ground = std::atoi(argv[4]) != 0;
            params.extrapolate_above_top = std::atoi(argv[5]) != 0;
            params.require_winds = true;
            const double dz = std::atof(argv[6]);
            std::vector<double> z;
            for (int a = 7; a < argc; ++a) z.push_back(std::atof(argv[a]));
            const tmv::init::FileSoundingSource source(params);
            const tmv::init::Sounding s = source.build(z, dz);
            s.verify_self_consistent();
            for (std::size_t k = 0; k < s.size(); ++k)
                std::printf("center %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", s.z_m[k], s.theta_k[k],
                            s.T_k[k], s.p_pa[k], s.rho_kgm3[k], s.qv_kgkg[k], s.u_ms[k], s.v_ms[k]);
            return 0;
        }
        return 2;
    }
    catch (const std::exception& e)
    {
        std::printf("error %s\n", e.what());
        return 0;
    }
}
