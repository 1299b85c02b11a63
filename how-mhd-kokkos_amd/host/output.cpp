// Output writers of the host shell: ASCII `dat` and binary appended-raw VTS
// StructuredGrid, restating src/output.cpp:21-571 byte for byte at one rank
// with the MPI build's file naming (output/dumpNNNNNN_rank000000.ext,
// output.cpp:21-37).  `output_format none` (a planned item of the reference,
// README.md:759) writes nothing.
#include "howmhd_host.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct View {
    const double* q;
    size_t n6y, n6z, ncell;
    double operator()(int m, int i, int j, int k) const {
        return q[(size_t)m * ncell + ((size_t)i * n6y + (size_t)j) * n6z + (size_t)k];
    }
};

std::string make_output_filename(const char* dir, int step, const std::string& ext) {
    std::ostringstream f;
    if (dir && dir[0]) f << dir << "/";
    f << "output/dump" << std::setw(6) << std::setfill('0') << step << "_rank" << std::setw(6)
      << std::setfill('0') << 0 << ext;
    return f.str();
}

// bc_index_host, src/output.cpp:43-59
int bc_index_host(int idx, int lo, int hi, bool periodic) {
    if (periodic) {
        const int n = hi - lo + 1;
        int r = (idx - lo) % n;
        if (r < 0) r += n;
        return lo + r;
    }
    if (idx < lo) return lo;
    if (idx > hi) return hi;
    return idx;
}

// Cell-centred diagnostic divB, src/output.cpp:154-188 and 314-347.
double cell_divb(const howmhd_run& par, const View& qh, int i, int j, int k) {
    const int ic_lo = 3, jc_lo = 3, kc_lo = 3;
    const int ic_hi = par.nx + 2, jc_hi = par.ny + 2, kc_hi = par.nz + 2;
    const bool xp = std::strcmp(par.x1bc, "periodic") == 0;
    const bool yp = std::strcmp(par.x2bc, "periodic") == 0;
    const bool zp = std::strcmp(par.x3bc, "periodic") == 0;
    double divb = 0.0;
    if (par.nx > 1) {
        const int im1 = bc_index_host(i - 1, ic_lo, ic_hi, xp), ip1 = bc_index_host(i + 1, ic_lo, ic_hi, xp);
        if (im1 == i || ip1 == i) divb += (qh(4, ip1, j, k) - qh(4, im1, j, k)) / par.dx;
        else divb += (qh(4, ip1, j, k) - qh(4, im1, j, k)) / (2.0 * par.dx);
    }
    if (par.ny > 1) {
        const int jm1 = bc_index_host(j - 1, jc_lo, jc_hi, yp), jp1 = bc_index_host(j + 1, jc_lo, jc_hi, yp);
        if (jm1 == j || jp1 == j) divb += (qh(5, i, jp1, k) - qh(5, i, jm1, k)) / par.dy;
        else divb += (qh(5, i, jp1, k) - qh(5, i, jm1, k)) / (2.0 * par.dy);
    }
    if (par.nz > 1) {
        const int km1 = bc_index_host(k - 1, kc_lo, kc_hi, zp), kp1 = bc_index_host(k + 1, kc_lo, kc_hi, zp);
        if (km1 == k || kp1 == k) divb += (qh(6, i, j, kp1) - qh(6, i, j, km1)) / par.dz;
        else divb += (qh(6, i, j, kp1) - qh(6, i, j, km1)) / (2.0 * par.dz);
    }
    return divb;
}

// output_dat, src/output.cpp:65-219
int output_dat(const howmhd_run& par, const View& qh, int step, double time, const char* dir) {
    const std::string filename = make_output_filename(dir, step, ".dat");
    std::filesystem::create_directories(std::filesystem::path(filename).parent_path());
    std::ofstream fout(filename);
    if (!fout) {
        std::cerr << "Error: cannot open output file " << filename << "\n";
        return 0;
    }
    fout << std::scientific << std::setprecision(15);
    fout << "# step = " << step << "\n";
    fout << "# time = " << time << "\n";
    fout << "# rank = " << 0 << "\n";
    fout << "# local_active_size = " << par.nx << " " << par.ny << " " << par.nz << "\n";
    fout << "# global_active_size = " << par.gnx << " " << par.gny << " " << par.gnz << "\n";
    fout << "# global_start_index = " << 0 << " " << 0 << " " << 0 << "\n";
    fout << "# columns: ig jg kg il jl kl x y z rho Mx My Mz Bx By Bz E pg vx vy vz divB\n";
    for (int k = 3; k <= par.nz + 2; ++k)
        for (int j = 3; j <= par.ny + 2; ++j)
            for (int i = 3; i <= par.nx + 2; ++i) {
                const int il = i - 3, jl = j - 3, kl = k - 3;
                const int ig = il, jg = jl, kg = kl;
                const double x = (static_cast<double>(ig) + 0.5) * par.dx;
                const double y = (static_cast<double>(jg) + 0.5) * par.dy;
                const double z = (static_cast<double>(kg) + 0.5) * par.dz;
                const double rho = qh(0, i, j, k), Mx = qh(1, i, j, k), My = qh(2, i, j, k), Mz = qh(3, i, j, k);
                const double Bx = qh(4, i, j, k), By = qh(5, i, j, k), Bz = qh(6, i, j, k), E = qh(7, i, j, k);
                const double vx = Mx / rho, vy = My / rho, vz = Mz / rho;
                const double vv2 = vx * vx + vy * vy + vz * vz;
                const double BB2 = Bx * Bx + By * By + Bz * Bz;
                const double pg = (par.gam - 1.0) * (E - 0.5 * rho * vv2 - 0.5 * BB2);
                const double divB = cell_divb(par, qh, i, j, k);
                fout << ig << " " << jg << " " << kg << " " << il << " " << jl << " " << kl << " " << x << " "
                     << y << " " << z << " " << rho << " " << Mx << " " << My << " " << Mz << " " << Bx << " "
                     << By << " " << Bz << " " << E << " " << pg << " " << vx << " " << vy << " " << vz << " "
                     << divB << "\n";
            }
    fout.close();
    std::cout << "Wrote " << filename << "\n";
    return 0;
}

// output_vts_binary, src/output.cpp:226-552
int output_vts(const howmhd_run& par, const View& qh, int step, double time, const char* dir) {
    const std::string filename = make_output_filename(dir, step, ".vts");
    std::filesystem::create_directories(std::filesystem::path(filename).parent_path());
    const int nx = par.nx, ny = par.ny, nz = par.nz;
    const size_t npts = (size_t)nx * ny * nz;
    std::vector<double> rho(npts), pressure(npts), energy(npts), v2(npts), B2(npts), divB(npts);
    std::vector<double> velocity(3 * npts), magnetic_field(3 * npts), momentum(3 * npts), points(3 * npts);
    for (int k = 3; k <= nz + 2; ++k)
        for (int j = 3; j <= ny + 2; ++j)
            for (int i = 3; i <= nx + 2; ++i) {
                const int ii = i - 3, jj = j - 3, kk = k - 3;
                const size_t p = (size_t)ii + (size_t)nx * ((size_t)jj + (size_t)ny * (size_t)kk);
                const double x = (static_cast<double>(ii) + 0.5) * par.dx;
                const double y = (static_cast<double>(jj) + 0.5) * par.dy;
                const double z = (static_cast<double>(kk) + 0.5) * par.dz;
                const double DD = qh(0, i, j, k), Mx = qh(1, i, j, k), My = qh(2, i, j, k), Mz = qh(3, i, j, k);
                const double Bx = qh(4, i, j, k), By = qh(5, i, j, k), Bz = qh(6, i, j, k), EE = qh(7, i, j, k);
                const double vx = Mx / DD, vy = My / DD, vz = Mz / DD;
                const double vv2 = vx * vx + vy * vy + vz * vz;
                const double BB2 = Bx * Bx + By * By + Bz * Bz;
                const double pg = (par.gam - 1.0) * (EE - 0.5 * DD * vv2 - 0.5 * BB2);
                rho[p] = DD; pressure[p] = pg; energy[p] = EE; v2[p] = vv2; B2[p] = BB2;
                divB[p] = cell_divb(par, qh, i, j, k);
                velocity[3 * p] = vx; velocity[3 * p + 1] = vy; velocity[3 * p + 2] = vz;
                magnetic_field[3 * p] = Bx; magnetic_field[3 * p + 1] = By; magnetic_field[3 * p + 2] = Bz;
                momentum[3 * p] = Mx; momentum[3 * p + 1] = My; momentum[3 * p + 2] = Mz;
                points[3 * p] = x; points[3 * p + 1] = y; points[3 * p + 2] = z;
            }
    std::uint64_t offset = 0;
    auto add_block = [&](std::uint64_t nbytes) {
        const std::uint64_t old = offset;
        offset += sizeof(std::uint64_t) + nbytes;
        return old;
    };
    const std::uint64_t off_rho = add_block(sizeof(double) * rho.size());
    const std::uint64_t off_pressure = add_block(sizeof(double) * pressure.size());
    const std::uint64_t off_energy = add_block(sizeof(double) * energy.size());
    const std::uint64_t off_v2 = add_block(sizeof(double) * v2.size());
    const std::uint64_t off_B2 = add_block(sizeof(double) * B2.size());
    const std::uint64_t off_divB = add_block(sizeof(double) * divB.size());
    const std::uint64_t off_velocity = add_block(sizeof(double) * velocity.size());
    const std::uint64_t off_magnetic_field = add_block(sizeof(double) * magnetic_field.size());
    const std::uint64_t off_momentum = add_block(sizeof(double) * momentum.size());
    const std::uint64_t off_points = add_block(sizeof(double) * points.size());

    std::ofstream fout(filename, std::ios::binary);
    if (!fout) {
        std::cerr << "Error: cannot open output file " << filename << "\n";
        return 0;
    }
    fout << "<?xml version=\"1.0\"?>\n";
    fout << "<VTKFile type=\"StructuredGrid\" version=\"0.1\" "
         << "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
    fout << "  <StructuredGrid WholeExtent=\"0 " << par.gnx - 1 << " 0 " << par.gny - 1 << " 0 " << par.gnz - 1
         << "\">\n";
    fout << "    <Piece Extent=\"" << 0 << " " << nx - 1 << " " << 0 << " " << ny - 1 << " " << 0 << " " << nz - 1
         << "\">\n";
    fout << "      <FieldData>\n";
    fout << "        <DataArray type=\"Float64\" Name=\"time\" NumberOfTuples=\"1\" format=\"ascii\">\n";
    fout << "          " << std::scientific << std::setprecision(15) << time << "\n";
    fout << "        </DataArray>\n";
    fout << "        <DataArray type=\"Int32\" Name=\"step\" NumberOfTuples=\"1\" format=\"ascii\">\n";
    fout << "          " << step << "\n";
    fout << "        </DataArray>\n";
    fout << "        <DataArray type=\"Int32\" Name=\"rank\" NumberOfTuples=\"1\" format=\"ascii\">\n";
    fout << "          " << 0 << "\n";
    fout << "        </DataArray>\n";
    fout << "        <DataArray type=\"Int32\" Name=\"global_start_index\" "
         << "NumberOfComponents=\"3\" NumberOfTuples=\"1\" format=\"ascii\">\n";
    fout << "          " << 0 << " " << 0 << " " << 0 << "\n";
    fout << "        </DataArray>\n";
    fout << "      </FieldData>\n";
    fout << "      <PointData>\n";
    auto arr = [&](const char* name, int ncomp, std::uint64_t off) {
        fout << "        <DataArray type=\"Float64\" Name=\"" << name << "\" "
             << "NumberOfComponents=\"" << ncomp << "\" format=\"appended\" offset=\"" << off << "\"/>\n";
    };
    arr("rho", 1, off_rho);
    arr("pressure", 1, off_pressure);
    arr("energy", 1, off_energy);
    arr("v2", 1, off_v2);
    arr("B2", 1, off_B2);
    arr("divB", 1, off_divB);
    arr("velocity", 3, off_velocity);
    arr("magnetic_field", 3, off_magnetic_field);
    arr("momentum", 3, off_momentum);
    fout << "      </PointData>\n";
    fout << "      <CellData>\n";
    fout << "      </CellData>\n";
    fout << "      <Points>\n";
    fout << "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"appended\" offset=\"" << off_points
         << "\"/>\n";
    fout << "      </Points>\n";
    fout << "    </Piece>\n";
    fout << "  </StructuredGrid>\n";
    fout << "  <AppendedData encoding=\"raw\">\n";
    fout << "_";
    auto write_block = [&](const std::vector<double>& data) {
        const std::uint64_t nbytes = static_cast<std::uint64_t>(sizeof(double) * data.size());
        fout.write(reinterpret_cast<const char*>(&nbytes), sizeof(std::uint64_t));
        fout.write(reinterpret_cast<const char*>(data.data()), static_cast<std::streamsize>(nbytes));
    };
    write_block(rho); write_block(pressure); write_block(energy); write_block(v2); write_block(B2);
    write_block(divB); write_block(velocity); write_block(magnetic_field); write_block(momentum);
    write_block(points);
    fout << "\n";
    fout << "  </AppendedData>\n";
    fout << "</VTKFile>\n";
    fout.close();
    std::cout << "Wrote " << filename << "  [binary VTS]\n";
    return 0;
}

}  // namespace

extern "C" int howmhd_host_output(const howmhd_run* run, const double* q0, int step, double time,
                                  const char* dir, char* errbuf, int errlen) {
    const size_t n6x = (size_t)run->nx + 6, n6y = (size_t)run->ny + 6, n6z = (size_t)run->nz + 6;
    View v{q0, n6y, n6z, n6x * n6y * n6z};
    const std::string fmt = run->output_format;
    if (fmt == "dat") return output_dat(*run, v, step, time, dir);
    if (fmt == "vtk" || fmt == "vts") return output_vts(*run, v, step, time, dir);
    if (fmt == "none") return 0;
    if (errbuf && errlen > 0)
        std::snprintf(errbuf, (size_t)errlen, "Error: unknown output_format = %s\nAvailable options: dat, vtk, none",
                      run->output_format);
    return HOWMHD_EINVAL;
}
