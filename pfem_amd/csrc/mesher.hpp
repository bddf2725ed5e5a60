// Host remesher (mesher.cpp): the node cloud, its alpha-shape triangulation and the remesh
// operations of the reference's Mesh class (srcs/mesh/Mesh.hpp).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace pfem {
namespace mesh {

struct MesherParams {
    int dim = 2;
    double hchar = 0, alpha = 1, omega = 1e16, gamma = 0;  // MeshCreateInfo defaults (Mesh.hpp:19-32)
    bool add_on_fs = true, delete_flying_nodes = false, laplacian_smoothing_boundaries = false;
    bool has_bounding_box = false;
    double bounding_box[6] = {0, 0, 0, 0, 0, 0};  // [min..., max...] over dim
    std::vector<std::array<double, 6>> exclusion_zones;
};

class Mesher {
public:
    // node flag bits (the PFEM_NODE_* values of the C-ABI)
    static constexpr uint8_t kBound = 1, kFree = 2, kOnFS = 4, kFixed = 8, kDirichlet = 16, kDirichletT = 32;

    explicit Mesher(const MesherParams& prm);

    // pos: dim x n, states: nstates x n (field-blocked); flags keep BOUND, FIXED, DIRICHLET
    void set_nodes(int n, int nstates, const double* pos, const double* states, const uint8_t* flags,
                   const int32_t* tags);
    // same node count as the last triangulation or set_nodes (positions / states after steps)
    void set_positions_states(const double* pos, const double* states);
    void triangulate();  // Mesh::triangulateAlphaShape
    void remesh();       // Mesh::remesh

    int n_nodes() const { return N_; }
    int n_states() const { return ns_; }
    int n_elems() const { return (int)(elems_.size() / (prm_.dim + 1)); }
    int n_facets() const { return (int)fout_.size(); }
    const MesherParams& params() const { return prm_; }
    void get_nodes(double* pos, double* states, uint8_t* flags, int32_t* tags) const;
    const std::vector<int>& elems() const { return elems_; }
    const std::vector<int>& facets() const { return facets_; }
    const std::vector<int>& facet_out() const { return fout_; }
    const std::vector<int>& facet_elem() const { return felem_; }

    // last remesh: nodes added / removed / out of the box / flying nodes deleted / duplicates
    // skipped; seconds in node operations, Delaunay, alpha shape and adjacency
    int added_ = 0, removed_ = 0, out_of_box_ = 0, flying_deleted_ = 0, skipped_ = 0;
    double t_nodes_ = 0, t_delaunay_ = 0, t_alpha_ = 0;

private:
    MesherParams prm_;
    int N_ = 0, ns_ = 0;
    std::vector<double> pos_, st_;  // node-major
    std::vector<uint8_t> flags_;
    std::vector<int32_t> tag_;
    std::vector<int> elems_, facets_, fout_, felem_;
    std::vector<int> nbr_ptr_, nbr_, nelem_;
    bool triangulated_ = false;

    bool is_onfs(int i) const;
    bool is_bound(int i) const;
    void build_adjacency();
    void erase_nodes(const std::vector<uint8_t>& del, bool renumber);
    void delete_flying_nodes();
    void laplacian_smoothing_boundaries();
    int add_nodes();
    int remove_nodes();
    int check_bounding_box();
};

}  // namespace mesh
}  // namespace pfem
