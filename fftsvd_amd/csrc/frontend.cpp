// DelPhi-style front end of the reference interface: parameter file,
// PDB / CRD / XYZR / SIZ / CRG readers, MODPDB / FRC writers, radius and
// charge assignment, and flat surface generation with this library's own
// SES mesher.  Host code only.  Same names, argument meaning, file formats
// and error behaviour as the reference (FFTSVDpbeAPI.c:36-672,1034-1049).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <strings.h>
#include <vector>

#include "../../include/FFTSVDpbeAPI.h"
#include "../../include/fftsvd_hip.h"

namespace {

// fgets-compatible line reader with the reference's 80-character lines
bool read_line(FILE* f, char* buf, int n) { return std::fgets(buf, n, f) != nullptr; }

// Copies the fixed-width field [pos, pos+len) of a line (shorter lines give
// a shorter or empty field) into out (NUL-terminated).
void field(const char* line, int pos, int len, char* out) {
   const int L = (int)std::strlen(line);
   int k = 0;
   for (int i = pos; i < pos + len && i < L; i++) out[k++] = line[i];
   out[k] = '\0';
}

char at(const char* line, int pos) { return pos < (int)std::strlen(line) ? line[pos] : '\0'; }

FILE* open_or_die(const char* what, const char* filename, const char* mode) {
   FILE* f = std::fopen(filename, mode);
   if (!f) error((char*)"Could not open %s file %s: %s", what, filename, std::strerror(errno));
   return f;
}

bool is_atom_line(const char* line) {
   return !strncasecmp(line, "ATOM", 4) || !strncasecmp(line, "HETATM", 6);
}

// Lines of a SIZ/CRG file after its first (title) line that are entries:
// every line not starting with '!' (an empty line still counts).
std::vector<std::string> table_lines(FILE* f) {
   std::vector<std::string> out;
   char line[81];
   if (!read_line(f, line, 81)) return out;
   while (read_line(f, line, 81))
      if (line[0] != '!' && std::strlen(line) > 0) out.emplace_back(line);
   return out;
}

}  // namespace

extern "C" {

const char* delphiparameterfilename = "fort.10";
const char* delphisizfilename = "fort.11";
const char* delphicrgfilename = "fort.12";
const char* delphipdbinputfilename = "fort.13";
const char* delphiphioutputfilename = "fort.14";
const char* delphipdbcollectfilename = "fort.15";
const char* delphifrcfilename = "fort.16";
const char* delphiphiinputfilename = "fort.18";
const char* delphipdboutputfilename = "fort.19";
const char* delphisrffilename = "fort.20";

// reference FFTSVDpbeAPI.c:64-79
void removeWhitespace(char* s) {
   char* w = s;
   for (char* r = s; *r; r++)
      if (!std::isspace((unsigned char)*r)) *w++ = *r;
   *w = '\0';
}

// reference FFTSVDpbeAPI.c:81-92
int yesno(char* s) {
   if (std::strlen(s) == 0) return 0;
   for (const char* t : {"t", "y", "true", "yes", "1"})
      if (!strcasecmp(s, t)) return 1;
   for (const char* f : {"f", "n", "false", "no", "0"})
      if (!strcasecmp(s, f)) return 0;
   error((char*)"Invalid parameter boolean expression");
   return 0;
}

// reference FFTSVDpbeAPI.c:94-169: "NAME=value" and "NAME(sub)" directives,
// whitespace ignored, names case-insensitive.
void readParams(const char* filename) {
   FILE* f = open_or_die("parameter", filename, "r");
   char line[81];
   while (read_line(f, line, 81)) {
      removeWhitespace(line);
      char command[25] = {0};
      int n = 0;
      while (n < 24 && std::isalpha((unsigned char)line[n])) {
         command[n] = (char)std::toupper((unsigned char)line[n]);
         n++;
      }
      char* value = line + n + (line[n] ? 1 : 0);   // after the '=' (or the '(')
      const std::string c(command);
      if (c == "DVERT") dielectricvertexdensity = std::atof(value);
      else if (c == "SVERT") saltvertexdensity = std::atof(value);
      else if (c == "INDI") innerdielectric = std::atof(value);
      else if (c == "EXDI") outerdielectric = std::atof(value);
      else if (c == "INFTYDI") inftydielectric = std::atof(value);
      else if (c == "LAMBDA") lambda = std::atof(value);
      else if (c == "SALT") ionicstrength = std::atof(value);
      else if (c == "IONRAD") ionexclusionradius = std::atof(value);
      else if (c == "PRBRAD") proberadius = std::atof(value);
      else if (c == "DE") errortolerance = std::atof(value);
      else if (c == "ADDCOUL") {
         if (yesno(value)) addcoulombic = 1;
      } else if (c == "USEQUAL") {
         if (yesno(value)) usequalocation = 1;
      } else if (c == "CURVED") {
         if (yesno(value)) usecurved = 1;
      } else if (c == "ENERGY") {
         // ENERGY(G): the character inside the parentheses
         const char* p = line + n;
         if (p[0] == '(' && p[1]) energycalculation = (char)std::toupper((unsigned char)p[1]);
      } else if (c == "OUT") {
         const char* p = line + n;
         std::string sub;
         if (p[0] == '(')
            for (p++; *p && *p != ')' && sub.size() < 9; p++) sub += *p;
         if (!strcasecmp(sub.c_str(), "MODPDB")) outputmodpdb = 1;
         else if (!strcasecmp(sub.c_str(), "FRC")) outputfrc = 1;
         else if (!strcasecmp(sub.c_str(), "PHI")) {
         } else warning((char*)"Unknown OUT parameter subcommand %s", sub.c_str());
      } else {
         warning((char*)"Unknown parameter command %s", command);
      }
   }
   std::fclose(f);
}

// reference FFTSVDpbeAPI.c:171-200
void checkParams() {
   if (dielectricvertexdensity < 0.0) error((char*)"Dielectric vertex density must be > 0.0");
   if (saltvertexdensity < 0.0) error((char*)"Salt vertex density must be > 0.0");
   if (innerdielectric < 1.0) error((char*)"Inner dielectric must be >= 1.0");
   if (outerdielectric < 1.0) error((char*)"Outer dielectric must be >= 1.0");
   if (innerdielectric > outerdielectric) warning((char*)"Inner dielectric is greater than outer dielectric");
   if (ionicstrength < 0.0) error((char*)"Ionic strength must be >= 0.0");
   if (ionexclusionradius < 0.0) error((char*)"Ionic exclusion radius must be >= 0.0");
   if (proberadius < 0.0) error((char*)"Probe radius must be >= 0.0");
   if (errortolerance <= 0.0) error((char*)"Error tolerance must be >= 0.0");
   if (energycalculation != 'G') error((char*)"Only energy calculation 'G' is currently supported");
   if (usequalocation && ionicstrength > 0.0) error((char*)"Qualocation does not work with ionic strength > 0.0");
}

// reference FFTSVDpbeAPI.c:202-276: ATOM/HETATM records, PDB fixed columns
void readPDB(const char* filename, unsigned int* numPDBentries, PDBentry** PDBentries) {
   FILE* f = open_or_die("PDB", filename, "r");
   char line[81], num[16];
   std::vector<PDBentry> v;
   while (read_line(f, line, 81)) {
      if (!is_atom_line(line)) continue;
      PDBentry e;
      std::memset(&e, 0, sizeof(e));
      field(line, 0, 6, e.record);
      field(line, 6, 5, num);
      e.atomnumber = (unsigned)std::atoi(num);
      field(line, 12, 4, e.atomname);
      e.alternatelocation = at(line, 16);
      field(line, 17, 3, e.residuename);
      e.chain = at(line, 21);
      field(line, 22, 4, num);
      e.residuenumber = (unsigned)std::atoi(num);
      e.residueinsertion = at(line, 26);
      field(line, 30, 8, num);
      e.x = std::atof(num);
      field(line, 38, 8, num);
      e.y = std::atof(num);
      field(line, 46, 8, num);
      e.z = std::atof(num);
      field(line, 54, 6, num);
      e.occupancy = std::atof(num);
      field(line, 60, 6, num);
      e.temperature = std::atof(num);
      field(line, 67, 3, num);
      e.footnotenumber = (unsigned)std::atof(num);
      removeWhitespace(e.record);
      removeWhitespace(e.atomname);
      removeWhitespace(e.residuename);
      v.push_back(e);
   }
   std::fclose(f);
   *numPDBentries = (unsigned)v.size();
   *PDBentries = (PDBentry*)std::calloc(v.size() + 1, sizeof(PDBentry));
   if (!*PDBentries) error((char*)"Error in memory allocation: %s", std::strerror(errno));
   std::copy(v.begin(), v.end(), *PDBentries);
}

// reference FFTSVDpbeAPI.c:278-335: CHARMM coordinates ('*' title lines,
// atom count, fixed columns)
void readCRD(const char* filename, unsigned int* numPDBentries, PDBentry** PDBentries) {
   FILE* f = open_or_die("CRD", filename, "r");
   char line[81], num[16];
   line[0] = '\0';
   while (read_line(f, line, 81))
      if (line[0] != '*') break;
   const int count = std::max(0, std::atoi(line));
   *PDBentries = (PDBentry*)std::calloc((size_t)count + 1, sizeof(PDBentry));
   if (!*PDBentries) error((char*)"Error in memory allocation: %s", std::strerror(errno));
   int k = 0;
   while (k < count && read_line(f, line, 81)) {
      PDBentry& e = (*PDBentries)[k];
      field(line, 0, 5, num);
      e.atomnumber = (unsigned)std::atoi(num);
      field(line, 5, 5, num);
      e.residuenumber = (unsigned)std::atoi(num);
      field(line, 11, 3, e.residuename);
      field(line, 16, 4, e.atomname);
      field(line, 20, 10, num);
      e.x = std::atof(num);
      field(line, 30, 10, num);
      e.y = std::atof(num);
      field(line, 40, 10, num);
      e.z = std::atof(num);
      e.chain = at(line, 51);
      field(line, 61, 10, num);
      e.temperature = std::atof(num);
      removeWhitespace(e.atomname);
      removeWhitespace(e.residuename);
      k++;
   }
   std::fclose(f);
   *numPDBentries = (unsigned)k;
}

// reference FFTSVDpbeAPI.c:337-372: "x y z r" per line
void readXYZR(const char* filename, unsigned int* numPDBentries, PDBentry** PDBentries) {
   FILE* f = open_or_die("XYZR", filename, "r");
   char line[81];
   std::vector<PDBentry> v;
   while (read_line(f, line, 81)) {
      PDBentry e;
      std::memset(&e, 0, sizeof(e));
      std::sscanf(line, "%lf %lf %lf %lf", &e.x, &e.y, &e.z, &e.radius);
      v.push_back(e);
   }
   std::fclose(f);
   *numPDBentries = (unsigned)v.size();
   *PDBentries = (PDBentry*)std::calloc(v.size() + 1, sizeof(PDBentry));
   std::copy(v.begin(), v.end(), *PDBentries);
}

// reference FFTSVDpbeAPI.c:374-434: DelPhi radius table
void readSIZ(const char* filename, unsigned int* numSIZentries, SIZentry** SIZentries) {
   FILE* f = open_or_die("SIZ", filename, "r");
   const std::vector<std::string> lines = table_lines(f);
   std::fclose(f);
   *numSIZentries = (unsigned)lines.size();
   *SIZentries = (SIZentry*)std::calloc(lines.size() + 1, sizeof(SIZentry));
   if (!*SIZentries) error((char*)"Error in memory allocation: %s", std::strerror(errno));
   char num[16];
   for (size_t k = 0; k < lines.size(); k++) {
      SIZentry& e = (*SIZentries)[k];
      const char* l = lines[k].c_str();
      field(l, 0, 6, e.atomlabel);
      field(l, 6, 3, e.residuelabel);
      field(l, 9, 8, num);
      e.radius = std::atof(num);
      removeWhitespace(e.atomlabel);
      removeWhitespace(e.residuelabel);
   }
}

// reference FFTSVDpbeAPI.c:436-500: DelPhi charge table
void readCRG(const char* filename, unsigned int* numCRGentries, CRGentry** CRGentries) {
   FILE* f = open_or_die("CRG", filename, "r");
   const std::vector<std::string> lines = table_lines(f);
   std::fclose(f);
   *numCRGentries = (unsigned)lines.size();
   *CRGentries = (CRGentry*)std::calloc(lines.size() + 1, sizeof(CRGentry));
   if (!*CRGentries) error((char*)"Error in memory allocation: %s", std::strerror(errno));
   char num[16];
   for (size_t k = 0; k < lines.size(); k++) {
      CRGentry& e = (*CRGentries)[k];
      const char* l = lines[k].c_str();
      field(l, 0, 6, e.atomlabel);
      field(l, 6, 3, e.residuelabel);
      field(l, 9, 4, num);
      e.residuenumber = (unsigned)std::atoi(num);
      e.chain = at(l, 13);
      field(l, 14, 8, num);
      e.charge = std::atof(num);
      removeWhitespace(e.atomlabel);
      removeWhitespace(e.residuelabel);
   }
}

// reference FFTSVDpbeAPI.c:504-527: PDB with radius (55-60) and charge (61-67)
void writeMODPDB(const char* filename, PDBentry* PDBentries, unsigned int numPDBentries) {
   FILE* f = open_or_die("PDB", filename, "w");
   std::fprintf(f, "HEADER output from FFTSVDpbe\n");
   std::fprintf(f, "HEADER atom radii in columns 55-60\n");
   std::fprintf(f, "HEADER atom charges in columns 61-67\n");
   for (unsigned a = 0; a < numPDBentries; a++) {
      const PDBentry& e = PDBentries[a];
      std::fprintf(f, "%-6s%5u %4s%1c%3s %1c%4u%1c   %8.3f%8.3f%8.3f%6.2f%7.3f\n", e.record, e.atomnumber,
                   e.atomname, e.alternatelocation, e.residuename, e.chain, e.residuenumber, e.residueinsertion,
                   e.x, e.y, e.z, e.radius, e.charge);
   }
   std::fclose(f);
}

// reference FFTSVDpbeAPI.c:529-565: DelPhi site-potential (FRC) file
void writeFRC(const char* filename, PDBentry* PDBentries, unsigned int numPDBentries) {
   FILE* f = open_or_die("FRC", filename, "w");
   std::fprintf(f, "FFTSVDpbe SITE POTENTIAL FILE\n");
   std::fprintf(f, "grid size,percent fill: %3d %8f\n", 0, 0.0);
   std::fprintf(f, "inner,outer dielectric: %8f %8f\n", innerdielectric, outerdielectric);
   std::fprintf(f, "ionic strength (M): %8f\n", ionicstrength);
   std::fprintf(f, "ion excl., probe radius: %8f %8f\n", ionexclusionradius, proberadius);
   std::fprintf(f, "linear, nolinear iterations: %8d %8d\n", 0, 0);
   std::fprintf(f, "boundary condition: %2d\n", 0);
   std::fprintf(f, "Data Output: COORDINATES CHARGE POTENTIALS FIELDS\n");
   std::fprintf(f, "title: FFTSVDpbe: BEM PBE Solver\n\n\n");
   std::fprintf(f, "    ATOM COORDINATES (X,Y,Z)     CHARGE   GRID PT.    GRID FIELDS: (Ex, Ey, Ez)\n");
   for (unsigned a = 0; a < numPDBentries; a++) {
      const PDBentry& e = PDBentries[a];
      std::fprintf(f, "%9.3f%9.3f%9.3f%8.4f%15.6f  %9.4f%9.4f%9.4f\n", e.x, e.y, e.z, e.charge, e.potential, 0.0, 0.0,
                   0.0);
   }
   std::fprintf(f, "total energy =   %6f  kt\n", 0.0);
   std::fclose(f);
}

// reference FFTSVDpbeAPI.c:569-672.  Radius: exact (atom, residue) match,
// else the longest-prefix-free atom-label match of a residue-less entry,
// else a one-letter element entry.  Charge: the most specific of eight
// (chain, residue number, residue name, atom name) match levels.
void assignRadiiCharges(PDBentry* PDBentries, unsigned int numPDBentries, SIZentry* SIZentries,
                        unsigned int numSIZentries, CRGentry* CRGentries, unsigned int numCRGentries) {
   for (unsigned a = 0; a < numPDBentries; a++) {
      PDBentry& p = PDBentries[a];
      int level = 0;
      for (unsigned r = 0; r < numSIZentries; r++) {
         const SIZentry& s = SIZentries[r];
         const bool nores = std::strlen(s.residuelabel) == 0;
         if (!std::strcmp(p.atomname, s.atomlabel) && !std::strcmp(p.residuename, s.residuelabel)) {
            p.radius = s.radius;
            level = 3;
            break;
         } else if (nores && !std::strncmp(p.atomname, s.atomlabel, std::strlen(s.atomlabel))) {
            if (level < 2) {
               p.radius = s.radius;
               level = 2;
            }
         } else if (nores && p.atomname[0] == s.atomlabel[0] && std::strlen(s.atomlabel) == 1) {
            if (level < 1) {
               p.radius = s.radius;
               level = 1;
            }
         }
      }
      if (level == 0) {
         warning((char*)"Atom %d %s %s %c has no radius entry, defaulting to 0.0", p.atomnumber, p.atomname,
                 p.residuename, p.chain);
         p.radius = 0.0;
      } else if (level == 1) {
         warning((char*)"Atom %d %s %s %c using default element radius of %f", p.atomnumber, p.atomname,
                 p.residuename, p.chain, p.radius);
      }
      level = 0;
      for (unsigned c = 0; c < numCRGentries; c++) {
         const CRGentry& q = CRGentries[c];
         const bool ch = q.chain == p.chain, rn = q.residuenumber == p.residuenumber;
         const bool res = !std::strcmp(p.residuename, q.residuelabel), atm = !std::strcmp(p.atomname, q.atomlabel);
         int l = 0;
         if (ch && rn && res && atm) l = 8;
         else if (rn && res && atm) l = 7;
         else if (ch && res && atm) l = 6;
         else if (res && atm) l = 5;
         else if (ch && rn && atm) l = 4;
         else if (rn && atm) l = 3;
         else if (ch && atm) l = 2;
         else if (atm) l = 1;
         if (l == 8) {
            p.charge = q.charge;
            level = 8;
            break;
         }
         if (l > level) {
            p.charge = q.charge;
            level = l;
         }
      }
      if (level == 0 && numCRGentries > 0) {
         warning((char*)"Atom %d %s %s %c has no charge entry, defaulting to 0.0", p.atomnumber, p.atomname,
                 p.residuename, p.chain);
         p.charge = 0.0;
      }
   }
}

// One connected component of a mesh: vertices, 0-based faces (MSMS
// orientation) and the signed volume they enclose.
namespace {
struct MeshComp {
   std::vector<double> V;
   std::vector<int> F;
   double vol = 0.0;
};

std::vector<MeshComp> split_components(const double* V, int nv, const int* F, const int* comp_faces, int ncomp) {
   std::vector<MeshComp> out(ncomp);
   int f0 = 0;
   for (int k = 0; k < ncomp; k++) {
      MeshComp& c = out[k];
      std::vector<int> remap(nv, -1);
      for (int q = f0; q < f0 + comp_faces[k]; q++)
         for (int s = 0; s < 3; s++) {
            const int v = F[3 * q + s];
            if (remap[v] < 0) {
               remap[v] = (int)(c.V.size() / 3);
               c.V.insert(c.V.end(), V + 3 * (size_t)v, V + 3 * (size_t)v + 3);
            }
            c.F.push_back(remap[v]);
         }
      f0 += comp_faces[k];
      for (size_t q = 0; q < c.F.size(); q += 3) {
         const double* a = &c.V[3 * (size_t)c.F[q]];
         const double* b = &c.V[3 * (size_t)c.F[q + 1]];
         const double* d = &c.V[3 * (size_t)c.F[q + 2]];
         c.vol += (a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                   a[2] * (b[0] * d[1] - b[1] * d[0])) / 6.0;
      }
   }
   return out;
}

// Winding number of the closed, positively oriented surface `s` about p:
// the solid angles of its faces (Van Oosterom and Strackee) over 4 pi;
// 1 inside, 0 outside.
double winding(const MeshComp& s, const double* p) {
   double w = 0.0;
   for (size_t q = 0; q < s.F.size(); q += 3) {
      double r[3][3], l[3];
      for (int k = 0; k < 3; k++) {
         for (int d = 0; d < 3; d++) r[k][d] = s.V[3 * (size_t)s.F[q + k] + d] - p[d];
         l[k] = std::sqrt(r[k][0] * r[k][0] + r[k][1] * r[k][1] + r[k][2] * r[k][2]);
      }
      const double det = r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) -
                         r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
                         r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
      auto dot = [&](int i, int j) { return r[i][0] * r[j][0] + r[i][1] * r[j][1] + r[i][2] * r[j][2]; };
      const double den = l[0] * l[1] * l[2] + dot(0, 1) * l[2] + dot(0, 2) * l[1] + dot(1, 2) * l[0];
      w += 2.0 * std::atan2(det, den);
   }
   return w / (4.0 * M_PI);
}

// Every vertex of `inner` inside `outer` (meshmaker's isInsideExhaustive);
// with `all` false, its first vertex only (isInside).
bool comp_inside(const MeshComp& inner, const MeshComp& outer, bool all) {
   const size_t nv = inner.V.size() / 3;
   for (size_t v = 0; v < (all ? nv : std::min<size_t>(nv, 1)); v++)
      if (!(winding(outer, &inner.V[3 * v]) > 0.5)) return false;
   return nv > 0;
}
}  // namespace

// Flat surfaces for a molecule with this library's SES mesher, in place of
// the reference's external meshmaker + MSMS run (FFTSVDpbeAPI.c:1034-1040;
// meshmaker.c:119-262,600-690): atoms from a PDB, CRD or XYZR file (by
// extension), radii from the SIZ table (XYZR carries its own), probe
// `proberadius`, about `dielectricvertexdensity` vertices per square
// angstrom.  As meshmaker does with MSMS's components, every connected
// component enclosing a positive volume is a dielectric surface and every
// other one a dielectric cavity; with salt and an ion-exclusion radius the
// grown-sphere surface is meshed too (salt surfaces and salt cavities).
// Parents as meshmaker assigns them: a dielectric the first salt surface
// containing it, a dielectric cavity the first dielectric containing all
// its vertices, a salt cavity the first dielectric cavity containing it;
// 999 where none does.  Writes <srf>.d<k>, .c<k>, .s<k>, .l<k> .vert/.face
// next to the .srf, and the .srf.
void generateFlatSRF(const char* pdbfilename, const char* sizfilename, const char* srffilename) {
   // with salt and an ion-exclusion radius (the Stern layer) the salt
   // surface is meshed too: the union of the atom spheres grown by
   // ionexclusionradius (meshmaker.c:105-133, there MSMS with a vanishing
   // probe), at the salt vertex density
   const bool stern = ionicstrength > 0.0 && ionexclusionradius > 1e-6;
   const std::string in(pdbfilename);
   auto ends = [&](const char* e) {
      const size_t n = std::strlen(e);
      return in.size() >= n && !strcasecmp(in.c_str() + in.size() - n, e);
   };
   unsigned n = 0;
   PDBentry* atoms = nullptr;
   if (ends(".xyzr")) {
      readXYZR(pdbfilename, &n, &atoms);
   } else {
      if (ends(".crd")) readCRD(pdbfilename, &n, &atoms);
      else readPDB(pdbfilename, &n, &atoms);
      unsigned ns = 0;
      SIZentry* siz = nullptr;
      readSIZ(sizfilename, &ns, &siz);
      assignRadiiCharges(atoms, n, siz, ns, nullptr, 0);
      std::free(siz);
   }
   std::vector<double> xyzr;
   for (unsigned a = 0; a < n; a++)
      if (atoms[a].radius > 0.0) {
         xyzr.push_back(atoms[a].x);
         xyzr.push_back(atoms[a].y);
         xyzr.push_back(atoms[a].z);
         xyzr.push_back(atoms[a].radius);
      }
   std::free(atoms);
   if (xyzr.empty()) error((char*)"no atom with a positive radius in %s", pdbfilename);
   const double probe = proberadius > 0.0 ? proberadius : 1.4;
   const std::string srf(srffilename);
   const size_t slash = srf.rfind('/');
   const std::string stem = slash == std::string::npos ? srf : srf.substr(slash + 1);
   // the components of one surface kind, outer surfaces then cavities; the
   // mesher's vertex density is about 4.5 / h^2 per square angstrom
   auto mesh = [&](double density, bool grown, std::vector<MeshComp>& surfs, std::vector<MeshComp>& cavs) {
      const double h = std::sqrt(4.5 / (density > 0.0 ? density : 1.0));
      double* V = nullptr;
      int *F = nullptr, *cf = nullptr;
      int nv = 0, nf = 0, nc = 0;
      const int rc = fftsvd_mesh_components(xyzr.data(), (int)(xyzr.size() / 4), grown ? ionexclusionradius : probe, h,
                                            grown ? 1 : 0, &V, &nv, &F, &nf, &cf, &nc);
      if (rc) error((char*)"surface generation failed for %s", pdbfilename);
      for (MeshComp& c : split_components(V, nv, F, cf, nc)) (c.vol > 0.0 ? surfs : cavs).push_back(std::move(c));
      fftsvd_free(V);
      fftsvd_free(F);
      fftsvd_free(cf);
      if (surfs.empty()) error((char*)"No surface with positive volume was created");
   };
   auto write = [&](const std::string& tag, const MeshComp& c) {
      const std::string base = srf + tag;
      FILE* fv = open_or_die("vert", (base + ".vert").c_str(), "w");
      for (size_t i = 0; i < c.V.size(); i += 3) std::fprintf(fv, "%.9f %.9f %.9f\n", c.V[i], c.V[i + 1], c.V[i + 2]);
      std::fclose(fv);
      FILE* ff = open_or_die("face", (base + ".face").c_str(), "w");
      for (size_t i = 0; i < c.F.size(); i += 3)
         std::fprintf(ff, "%d %d %d 1\n", c.F[i] + 1, c.F[i + 1] + 1, c.F[i + 2] + 1);
      std::fclose(ff);
      return stem + tag;
   };
   std::vector<MeshComp> diel, dcav, salt, scav;
   mesh(dielectricvertexdensity, false, diel, dcav);
   if (stern) mesh(saltvertexdensity, true, salt, scav);
   // a cavity, oriented into its void, encloses the void with the opposite
   // sign: flip it for the containment tests of the salt cavities
   std::vector<MeshComp> dcav_pos = dcav;
   for (MeshComp& c : dcav_pos)
      for (size_t q = 0; q < c.F.size(); q += 3) std::swap(c.F[q + 1], c.F[q + 2]);
   auto names = [&](const char* kind, const std::vector<MeshComp>& cs) {
      std::string line;
      for (size_t k = 0; k < cs.size(); k++) line += (k ? " " : "") + write("." + std::string(kind) + std::to_string(k), cs[k]);
      return line;
   };
   auto parent = [](const MeshComp& c, const std::vector<MeshComp>& outers, bool all) {
      for (size_t k = 0; k < outers.size(); k++)
         if (comp_inside(c, outers[k], all)) return std::to_string(k);
      return std::string("999");
   };
   auto parents = [&](const std::vector<MeshComp>& cs, const std::vector<MeshComp>& outers, bool all) {
      std::string line;
      for (size_t k = 0; k < cs.size(); k++) line += (k ? " " : "") + parent(cs[k], outers, all);
      return line;
   };
   const std::string sline = names("s", salt), dline = names("d", diel), cline = names("c", dcav),
                     lline = names("l", scav);
   // .srf: panel types, salt surfaces, dielectrics, dielectric cavities,
   // salt cavities, then the parents of dielectrics and of both cavity kinds
   FILE* fs = open_or_die("surface", srffilename, "w");
   std::fprintf(fs, "f\nf\n%s\n%s\n%s\n%s\n%s\n%s\n%s\n", sline.c_str(), dline.c_str(), cline.c_str(), lline.c_str(),
                parents(diel, salt, false).c_str(), parents(dcav, diel, true).c_str(),
                parents(scav, dcav_pos, false).c_str());
   std::fclose(fs);
}

// reference FFTSVDpbeAPI.c:1042-1049 (curved panels are not supported here)
void generateCurvedSRF(const char* pdbfilename, const char* sizfilename, const char* srffilename) {
   (void)pdbfilename;
   (void)sizfilename;
   (void)srffilename;
   error((char*)"curved panels are not supported (flat surfaces only)");
}

}  // extern "C"
