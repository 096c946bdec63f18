// survey.hpp -- drop-in replacement of the reference's survey helper
// (include/survey.hpp:10-35): cosmology struct, (RA, Dec, z) -> Cartesian
// conversion and the box/origin preparation of survey catalogues.
#ifndef SURVEY_HPP
#define SURVEY_HPP

#include <vector>

#include "nc_vector_types.hpp"

struct cosmology {
    double H0;
    double OmegaM;
    double OmegaLambda;
};

class survey {
   public:
    survey(const cosmology &cosmo);

    // (RA deg, Dec deg, z, w) -> (x, y, z, w) in Mpc/h (survey.cu:39-57).
    std::vector<double4> convert_coordinates(const std::vector<double4> &catalog);

    // box = extent + 1.0, ori = -min, catalogues shifted by -min (survey.cu:106-128).
    void prep_cat_survey(std::vector<double4> &catalog, double3 &box, double3 &ori);
    void prep_cat_survey(std::vector<double4> &catalog1, std::vector<double4> &catalog2, double3 &box, double3 &ori);

    // Comoving distance in Mpc/h: c/100 * int_0^z dz'/sqrt(Om (1+z')^3 + OL).
    double comoving_distance(double z);

   private:
    void find_min_max_elements(const std::vector<double4> &data, double3 &catMin, double3 &catMax);
    void shift_catalog(std::vector<double4> &catalog, double3 catMin);

    cosmology cosmo_;
};

#endif  // SURVEY_HPP
