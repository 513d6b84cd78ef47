// C ABI of the solution output (include/cfdg.h, cfdg_write_solution_vtu): the
// solution files of one rank as the reference writes them -- a VTU piece with
// the mesh, a rank array and the cell fields rho, u, v, w, pressure, mach
// (with the turbulence module also nu_tilde and mu_t, appended by the module,
// spalart_allmaras.hpp:445-448), and its PVTU master (Solver::write_fields,
// solver.hpp:856-898; io::vtk::write_solution_vtu, vtu.cpp:20-81, 321-475).  The fields are formed
// on the device from the state and written in the mesh's cell order with the
// reference's text formats, so the files are byte-identical for equal states.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <filesystem>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "cfdg.h"
#include "ctx.hpp"
#include "kernels.hpp"

using cfdg::CfdgError;

namespace {

std::size_t z(std::int64_t v) { return static_cast<std::size_t>(v); }

// VTK cell type of a cell kind (vtu.cpp:20-30)
unsigned vtk_type(std::uint8_t kind) {
    switch (static_cast<cfdg::CellKind>(kind)) {
        case cfdg::CellKind::Tri: return 5;
        case cfdg::CellKind::Quad: return 9;
        case cfdg::CellKind::Tet: return 10;
        case cfdg::CellKind::Hexa: return 12;
        case cfdg::CellKind::Prism: return 13;
        case cfdg::CellKind::Pyra: return 14;
        default: return 0;
    }
}

struct File {
    FILE* f = nullptr;
    std::vector<char> buf = std::vector<char>(1 << 20);
    explicit File(const std::string& path) {
        f = std::fopen(path.c_str(), "wb");
        if (!f) throw CfdgError(CFDG_ERR_ARG, "cannot open " + path);
        std::setvbuf(f, buf.data(), _IOFBF, buf.size());
    }
    ~File() {
        if (f) std::fclose(f);
    }
};

const char* const kFieldNames[8] = {"rho", "u", "v", "w", "pressure", "mach", "nu_tilde", "mu_t"};

// fields: nf planes of n_cells in mesh order, named kFieldNames[0 .. nf)
void write_piece(const std::string& path, const cfdg::Mesh& m, const std::vector<double>& fields, int nf) {
    File out(path);
    FILE* f = out.f;
    const int nc = m.n_cells, np = m.n_nodes;
    std::fprintf(f, "<?xml version=\"1.0\"?>\n");
    std::fprintf(f, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    std::fprintf(f, "  <UnstructuredGrid>\n");
    std::fprintf(f, "    <Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\">\n", np, nc);
    std::fprintf(f, "      <Points>\n");
    std::fprintf(f, "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
    for (int i = 0; i < np; ++i)
        std::fprintf(f, "          %.16g %.16g %.16g\n", m.node_x[z(i)], m.node_y[z(i)], m.node_z[z(i)]);
    std::fprintf(f, "        </DataArray>\n");
    std::fprintf(f, "      </Points>\n");
    std::fprintf(f, "      <Cells>\n");
    std::fprintf(f, "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
    for (int c = 0; c < nc; ++c) {
        std::fprintf(f, "          ");
        for (std::int32_t k = m.cell_nodes_off[z(c)]; k < m.cell_nodes_off[z(c) + 1]; ++k)
            std::fprintf(f, "%d ", m.cell_nodes[z(k)]);
        std::fprintf(f, "\n");
    }
    std::fprintf(f, "        </DataArray>\n");
    std::fprintf(f, "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n          ");
    for (int c = 0; c < nc; ++c) std::fprintf(f, "%d ", m.cell_nodes_off[z(c) + 1]);
    std::fprintf(f, "\n        </DataArray>\n");
    std::fprintf(f, "        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n          ");
    for (int c = 0; c < nc; ++c) std::fprintf(f, "%u ", vtk_type(m.cell_type[z(c)]));
    std::fprintf(f, "\n        </DataArray>\n");
    std::fprintf(f, "      </Cells>\n");
    std::fprintf(f, "      <CellData>\n");
    std::fprintf(f, "        <DataArray type=\"Int32\" Name=\"rank\" format=\"ascii\">\n          ");
    for (int c = 0; c < nc; ++c) std::fprintf(f, "%d ", 0);
    std::fprintf(f, "\n        </DataArray>\n");
    for (int i = 0; i < nf; ++i) {
        std::fprintf(f, "        <DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n          ", kFieldNames[i]);
        const double* v = fields.data() + z(i) * z(nc);
        for (int c = 0; c < nc; ++c) std::fprintf(f, "%.10e ", v[z(c)]);
        std::fprintf(f, "\n        </DataArray>\n");
    }
    std::fprintf(f, "      </CellData>\n");
    std::fprintf(f, "    </Piece>\n");
    std::fprintf(f, "  </UnstructuredGrid>\n");
    std::fprintf(f, "</VTKFile>\n");
}

void write_master(const std::string& path, const std::string& stem, int nf) {
    File out(path);
    FILE* f = out.f;
    std::fprintf(f, "<?xml version=\"1.0\"?>\n");
    std::fprintf(f, "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    std::fprintf(f, "  <PUnstructuredGrid GhostLevel=\"0\">\n");
    std::fprintf(f, "    <PCellData>\n");
    std::fprintf(f, "      <PDataArray type=\"Int32\" Name=\"rank\"/>\n");
    for (int i = 0; i < nf; ++i) std::fprintf(f, "      <PDataArray type=\"Float64\" Name=\"%s\"/>\n", kFieldNames[i]);
    std::fprintf(f, "    </PCellData>\n");
    std::fprintf(f, "    <PPoints>\n");
    std::fprintf(f, "      <PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n");
    std::fprintf(f, "    </PPoints>\n");
    std::fprintf(f, "    <Piece Source=\"%s_%05d.vtu\"/>\n", stem.c_str(), 0);
    std::fprintf(f, "  </PUnstructuredGrid>\n");
    std::fprintf(f, "</VTKFile>\n");
}

}  // namespace

extern "C" int cfdg_write_solution_vtu(cfdg_ctx* c, const cfdg_mesh* mesh, const char* outdir, const char* stem) {
    if (!c) return CFDG_ERR_ARG;
    try {
        if (!mesh || !outdir || !stem) throw CfdgError(CFDG_ERR_ARG, "null argument");
        const cfdg::Mesh& m = mesh->m;
        if (m.n_cells != c->nc) throw CfdgError(CFDG_ERR_ARG, "mesh does not match the context");
        const std::size_t nc = z(c->nc);
        cfdg::DBuf<double> d;
        d.alloc(6 * nc);
        cfdg::launch_output_fields(c->nc, c->nargs.gas, c->U.get(), d.get(), c->stream);
        c->check_launch();
        // the module's nu_tilde and mu_t as of the last residual evaluation
        // (pre_sweep, spalart_allmaras.hpp:263-281), zero before the first
        const int nf = c->sa_on() ? 8 : 6;
        std::vector<double> dev(z(nf) * nc, 0.0), fields(z(nf) * nc);
        c->d2h(dev.data(), d.get(), 6 * nc);
        if (c->sa_on() && c->sa_evaluated) {
            c->d2h(dev.data() + 6 * nc, c->sa_nu.get(), nc);
            c->d2h(dev.data() + 7 * nc, c->sa_mut.get(), nc);
        }
        c->sync();
        for (std::size_t i = 0; i < z(nf); ++i)
            for (std::size_t cc = 0; cc < nc; ++cc) fields[i * nc + cc] = dev[i * nc + z(c->old2new[cc])];
        std::error_code ec;
        std::filesystem::create_directories(outdir, ec);
        const std::string base = std::string(outdir) + "/" + stem;
        write_piece(base + "_00000.vtu", m, fields, nf);
        write_master(base + ".pvtu", stem, nf);
        return CFDG_OK;
    } catch (const CfdgError& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return CFDG_ERR_ARG;
    }
}
