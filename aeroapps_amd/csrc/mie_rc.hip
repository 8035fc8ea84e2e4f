// Chem_Mie construction from resource files: the i90 reads of
// Chem_MieCreateFromRC (src/Shared/GMAO_Shared/Chem_Base/Chem_MieMod.F90:
// 124-411) and of the chemistry registry it maps the tables to
// (Chem_RegistryCreate, src/Shared/GMAO_Shared/Chem_Base/Chem_RegistryMod.F90:
// 200-545).  The optics-table files themselves are read by a caller-supplied
// reader (the reference reads netCDF with Chem_MieTableRead); their arrays
// then go through aa_mie_add_table_native, which does the rest of
// Chem_MieTableRead.  Host code only.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/aeroapps_amd.h"

namespace {

// Resource file in the i90 format of GMAO_mpeu's m_inpak90 as the two
// constructors use it: "label: token token ..." lines, '#' comments to the end
// of a line, and tables "label::" ... "::" whose rows are the lines between.
struct RcFile {
  std::vector<std::vector<std::string>> lines;   // tokens of each non-empty line

  bool load(const char* path) {
    std::ifstream f(path);
    if (!f) return false;
    std::string raw;
    while (std::getline(f, raw)) {
      const size_t h = raw.find('#');
      if (h != std::string::npos) raw.resize(h);
      for (char& c : raw) if (c == ',' || c == '\t' || c == '\r') c = ' ';
      std::istringstream is(raw);
      std::vector<std::string> tok;
      std::string w;
      while (is >> w) tok.push_back(w);
      lines.push_back(tok);
    }
    return true;
  }
  // Tokens after `label` on the line where it first appears ("label:" may be
  // written with its value attached, as i90 allows).  False when absent.
  bool label(const std::string& lab, std::vector<std::string>& rest) const {
    for (const auto& l : lines) {
      for (size_t i = 0; i < l.size(); ++i) {
        if (l[i] == lab) { rest.assign(l.begin() + i + 1, l.end()); return true; }
        if (l[i].compare(0, lab.size(), lab) == 0 && l[i].size() > lab.size()) {
          rest.assign(1, l[i].substr(lab.size()));
          rest.insert(rest.end(), l.begin() + i + 1, l.end());
          return true;
        }
      }
    }
    return false;
  }
  // Rows of a "label::" table (first token of each row), false when absent.
  bool table(const std::string& lab, std::vector<std::vector<std::string>>& rows) const {
    for (size_t k = 0; k < lines.size(); ++k) {
      if (lines[k].empty() || lines[k][0] != lab) continue;
      rows.clear();
      for (size_t j = k + 1; j < lines.size(); ++j) {
        if (!lines[j].empty() && lines[j][0] == "::") return true;
        if (!lines[j].empty()) rows.push_back(lines[j]);
      }
      return true;
    }
    return false;
  }
};

std::string upper(std::string s) {
  for (char& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

// Species the registry knows, in its index order (setidx_ calls), with the
// default name prefix of setmeta_.  The seven aerosol species with optics
// tables carry their AA_SP_* code.
struct Species { const char* name; const char* prefix; int code; };
const Species kSpecies[] = {
    {"H2O", "q", -1}, {"O3", "o3", -1}, {"CO", "CO", -1}, {"CO2", "CO2", -1}, {"DU", "du", AA_SP_DU},
    {"SS", "ss", AA_SP_SS}, {"SU", "su", AA_SP_SU}, {"CFC", "CFC", -1}, {"BC", "bc", AA_SP_BC},
    {"OC", "oc", AA_SP_OC}, {"BRC", "brc", AA_SP_BRC}, {"Rn", "Rn", -1}, {"CH4", "CH4", -1},
    {"NI", "ni", AA_SP_NI}, {"SC", "sc", -1}, {"GMI", "GMI", -1}, {"XX", "xx", -1}, {"OCS", "ocs", -1},
    {"TR", "TR", -1}};

void put_name(char* dst, const std::string& n) {
  for (int c = 0; c < 16; ++c) dst[c] = c < (int)n.size() ? n[c] : ' ';
}

}  // namespace

extern "C" {

int aa_mie_registry(const char* path, int nq_max, int* nq, char* names, int* species, int* rc) {
  if (!path || !nq || nq_max < 0 || !rc) return AA_EARG;
  *rc = 0;
  *nq = 0;
  RcFile f;
  if (!f.load(path)) { *rc = 1; return 0; }
  struct Entry { std::string name; int code; };
  std::vector<Entry> all;
  for (const Species& sp : kSpecies) {
    // parserc_: doing_XX: y/Y..., then nbins_XX: and units_XX: (required when doing)
    std::vector<std::string> v;
    bool doing = f.label(std::string("doing_") + sp.name + ":", v) && !v.empty() &&
                 (v[0][0] == 'y' || v[0][0] == 'Y');
    if (!doing) continue;
    if (!f.label(std::string("nbins_") + sp.name + ":", v) || v.empty()) { *rc = 2; return 0; }
    const int nb = std::atoi(v[0].c_str());
    std::vector<std::string> u;   // units_XX: required too ("cannot determine UNITS")
    if (!f.label(std::string("units_") + sp.name + ":", u) || u.empty()) { *rc = 2; return 0; }
    if (nb < 1) continue;   // setidx_: no tracers
    // setmeta_: names from variable_table_XX:: when present, else prefix + bin number
    std::vector<std::vector<std::string>> rows;
    const bool tab = f.table(std::string("variable_table_") + upper(sp.prefix) + "::", rows);
    for (int b = 0; b < nb; ++b) {
      std::string n;
      if (tab) {
        if (b >= (int)rows.size()) { *rc = 3; return 0; }   // "Not enough names in chemistry registry file"
        n = rows[b][0];
      } else {
        char buf[8];
        std::snprintf(buf, sizeof buf, "%03d", b + 1);
        n = std::string(sp.prefix) + buf;
      }
      all.push_back({n, sp.code});
    }
  }
  *nq = (int)all.size();
  if ((int)all.size() > nq_max) return 0;   // caller asks again with room for *nq names
  for (size_t i = 0; i < all.size(); ++i) {
    if (names) put_name(names + 16 * i, all[i].name);
    if (species) species[i] = all[i].code;
  }
  return 0;
}

int aa_mie_create_rc(aa_mie** out, const char* rcfile, int nq, const char* names, const int* species,
                     aa_mie_table_reader reader, void* user, int* rc) {
  if (!out || !rcfile || nq < 0 || (nq > 0 && (!names || !species)) || !reader || !rc) return AA_EARG;
  *out = nullptr;
  *rc = 0;
  RcFile f;
  if (!f.load(rcfile)) { *rc = 1; return 0; }
  std::vector<std::string> v;
  // n_channels (required), n_moments (0 when absent), r_channels [m]
  if (!f.label("n_channels:", v) || v.empty()) { *rc = 1; return 0; }
  const int nch = std::atoi(v[0].c_str());
  int nmom = 0;
  if (f.label("n_moments:", v) && !v.empty()) nmom = std::atoi(v[0].c_str());
  if (nch <= 0 || !f.label("r_channels:", v) || (int)v.size() < nch) { *rc = 1; return 0; }
  std::vector<float> ch(nch);
  for (int c = 0; c < nch; ++c) ch[c] = std::strtof(v[c].c_str(), nullptr);
  // the seven table files, read in the reference's order
  const int order[7] = {AA_SP_DU, AA_SP_SS, AA_SP_SU, AA_SP_OC, AA_SP_BC, AA_SP_NI, AA_SP_BRC};
  const char* label[7] = {"DU", "SS", "SU", "OC", "BC", "NI", "BRC"};
  std::string file[7];
  {
    const char* lab_of[7] = {"DU", "SS", "BC", "OC", "SU", "NI", "BRC"};   // the labels' order in the file
    for (const char* l : lab_of) {
      if (!f.label(std::string("filename_optical_properties_") + l + ":", v) || v.empty()) { *rc = 1; return 0; }
      for (int k = 0; k < 7; ++k) if (!std::strcmp(label[k], l)) file[k] = v[0];
    }
  }
  aa_mie* t = nullptr;
  int e = aa_mie_create(&t, nch, ch.data());
  if (e) return e;
  for (int k = 0; k < 7; ++k) {
    aa_mie_table_file tf;
    std::memset(&tf, 0, sizeof tf);
    if (reader(user, file[k].c_str(), &tf) != 0 || tf.nrh <= 0 || tf.nbin <= 0 || tf.nch <= 0) {
      aa_mie_destroy(t);
      *rc = 2;
      return 0;
    }
    // bins served: the registry's tracers of this species, in order (sulfate
    // only so4/SO4/sul/SUL names, nitrate only no3/NO3,
    // Chem_MieMod.F90:376-400); bins beyond them are served to no tracer
    std::vector<char> bn((size_t)16 * tf.nbin, ' ');
    int b = 0;
    for (int iq = 0; iq < nq && b < tf.nbin; ++iq) {
      if (species[iq] != order[k]) continue;
      std::string n(names + 16 * iq, 16);
      while (!n.empty() && (n.back() == ' ' || n.back() == '\0')) n.pop_back();
      const std::string p3 = n.substr(0, 3);
      if (order[k] == AA_SP_SU && !(p3 == "so4" || p3 == "SO4" || p3 == "sul" || p3 == "SUL")) continue;
      if (order[k] == AA_SP_NI && !(p3 == "no3" || p3 == "NO3")) continue;
      put_name(bn.data() + 16 * b, n);
      ++b;
    }
    int trc = 0;
    e = aa_mie_add_table_native(t, tf.nch, tf.lambda, tf.nrh, tf.nbin, tf.nmom, tf.npol, nmom, tf.rh, tf.reff,
                                tf.bext, tf.bsca, tf.bbck, tf.g, tf.pback, tf.pmom, tf.gf, tf.rhop, tf.refr,
                                tf.refi, bn.data(), &trc);
    if (e || trc) {
      aa_mie_destroy(t);
      if (e) return e;
      *rc = trc;
      return 0;
    }
  }
  *out = t;
  return 0;
}

}  // extern "C"
