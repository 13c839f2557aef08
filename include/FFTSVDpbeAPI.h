/*
 * Drop-in C interface of the MI355X FFT-SVD Poisson-Boltzmann BEM solver.
 *
 * Same function names, argument lists, value types and call protocol as
 * the reference's FFTSVDpbeAPI.h (reference FFTSVDpbeAPI.h:1-163) for the
 * solve path, so a program written against the reference (its FFTSVDpbe.c
 * driver, for instance) compiles against this header and links with
 * libfftsvd_hip.so.  Behind it, the operator, preconditioner, GMRES, right-
 * hand side and potential collection run on the GPU; host-visible vectors
 * are plain calloc'ed double arrays the caller frees with free(), as in the
 * reference.
 *
 * Types the caller dereferences keep the reference's field names and types
 * (PDBentry, Charge, the SurfaceOperator fields children / numchildren /
 * charges / mynumpanels / startphi / startdphi / epsilon / kappa); the rest
 * is opaque.
 */
#ifndef FFTSVDpbeAPI_H
#define FFTSVDpbeAPI_H

#ifdef __cplusplus
extern "C" {
#endif

#define VERSION "0.01"
#define KT_CONVERSION 561.0   /* reference FFTSVDpbeAPI.h:13 */
#define DEBYE_CONSTANT 3.047  /* reference FFTSVDpbeAPI.h:14 */
#define GRID_SIZE 4           /* reference FFTSVDpbeAPI.h:16 */
#define SVD_ERROR 1e-5        /* reference FFTSVDpbeAPI.h:17 */
#define MAX_PANELS_PER_FINEST_CUBE 32 /* reference FFTSVDpbeAPI.h:19 */

typedef double real;
typedef real* Vector;

typedef struct _Vector3D {
   real x, y, z;
} _Vector3D;
typedef _Vector3D* Vector3D;

typedef enum { CONSTANT_KERNEL, X_KERNEL, Y_KERNEL, Z_KERNEL, POISSON_KERNEL, HELMHOLTZ_KERNEL } BEMKernelType;

/* A flat panel as read from .vert/.face (reference Panel.h:3-27: centroid,
 * normal, area; the vertices are this library's addition). */
typedef struct _Panel {
   Vector3D centroid;
   Vector3D normal;
   real area;
   real maxedgelength;
   real vertex[3][3];
} _Panel;
typedef _Panel* Panel;

/* reference Charge.h:3-8 */
typedef struct _Charge {
   unsigned int numcharges;
   unsigned int globalindexstart;
   Vector3D* points;
   Vector charges;
} _Charge;
typedef _Charge* Charge;

/* reference SurfaceOperator.h:5-23; `tree` points at the library's device
 * operator for this region; M3 is non-NULL where an M3 tree was built
 * (accelerateM3 set before generateSurfaceOperator; the potentials of
 * collectPotentials then go through it, SurfaceOperator.c:347-363). */
typedef struct _SurfaceOperator {
   BEMKernelType kernel;
   real epsilon;
   real kappa;
   unsigned int mynumpanels;
   unsigned int startphi;
   unsigned int startdphi;
   void* tree;
   void* M3;
   unsigned int numchildren;
   struct _SurfaceOperator** children;
   Vector resultInternal;
   Vector resultExternal;
   struct _SurfaceOperator* parent;
   Charge charges;
} _SurfaceOperator;
typedef _SurfaceOperator* SurfaceOperator;

/* reference QualocationOperator.h:3-10; `tree` points at the library's
 * device operator; M1M3 is non-NULL when accelerateM3 built that tree. */
typedef struct _QualocationOperator {
   void* tree;
   void* M1M3;
   real innerdielectric, outerdielectric;
   Charge charges;
} _QualocationOperator;
typedef _QualocationOperator* QualocationOperator;

/* reference Preconditioner.h:3-17 (factors live on the device) */
typedef struct _Preconditioner {
   unsigned int rows, columns, nnz;
   void* device;
} _Preconditioner;
typedef _Preconditioner* Preconditioner;

/* reference FFTSVDpbeAPI.h:23-56 */
typedef struct {
   char record[7];
   unsigned int atomnumber;
   char atomname[5];
   char alternatelocation;
   char residuename[4];
   char chain;
   unsigned int residuenumber;
   char residueinsertion;
   real x;
   real y;
   real z;
   real occupancy;
   real temperature;
   unsigned int footnotenumber;
   real radius;
   real charge;
   real potential;
} PDBentry;

typedef struct {
   char atomlabel[7];
   char residuelabel[4];
   real radius;
} SIZentry;

typedef struct {
   char atomlabel[7];
   char residuelabel[4];
   unsigned int residuenumber;
   char chain;
   real charge;
} CRGentry;

/* DelPhi unit-number file names (reference FFTSVDpbeAPI.h:60-69, defaults
 * "fort.10" .. "fort.20", FFTSVDpbeAPI.c:5-14) */
extern const char* delphiparameterfilename;
extern const char* delphisizfilename;
extern const char* delphicrgfilename;
extern const char* delphipdbinputfilename;
extern const char* delphiphioutputfilename;
extern const char* delphipdbcollectfilename;
extern const char* delphifrcfilename;
extern const char* delphiphiinputfilename;
extern const char* delphipdboutputfilename;
extern const char* delphisrffilename;

/* Parameters (reference FFTSVDpbeAPI.h:70-89, defaults FFTSVDpbeAPI.c:18-34) */
extern real dielectricvertexdensity;
extern real saltvertexdensity;
extern real innerdielectric;
extern real outerdielectric;
extern real lambda;
extern real inftydielectric;
extern real ionicstrength;
extern real ionexclusionradius;
extern real proberadius;
extern real errortolerance;
extern char energycalculation;
extern unsigned int outputmodpdb;
extern unsigned int outputfrc;
extern unsigned int addcoulombic;
extern unsigned int accelerateM3;
extern unsigned int usequalocation;
extern unsigned int usecurved;

/* Utility (reference FFTSVDpbeAPI.h:93-96; FFTSVDpbeAPI.c:38-93) */
void error(char* message, ...);
void warning(char* message, ...);

void removeWhitespace(char* s);   /* reference FFTSVDpbeAPI.h:94; FFTSVDpbeAPI.c:64-79 */
int yesno(char* s);               /* reference FFTSVDpbeAPI.h:95; FFTSVDpbeAPI.c:81-92 */

/* Input (reference FFTSVDpbeAPI.h:99-105; FFTSVDpbeAPI.c:94-500): the
 * DelPhi-style parameter file (NAME=value, ENERGY(G), OUT(FRC|MODPDB|PHI)),
 * PDB / CHARMM CRD / XYZR coordinates and SIZ / CRG tables.  Arrays are
 * calloc'ed; the caller frees them. */
void readParams(const char* filename);
void checkParams();
void readPDB(const char* filename, unsigned int* numPDBentries, PDBentry** PDBentries);
void readCRD(const char* filename, unsigned int* numPDBentries, PDBentry** PDBentries);
void readXYZR(const char* filename, unsigned int* numPDBentries, PDBentry** PDBentries);
void readSIZ(const char* filename, unsigned int* numSIZentries, SIZentry** SIZentries);
void readCRG(const char* filename, unsigned int* numCRGentries, CRGentry** CRGentries);

/* Output (reference FFTSVDpbeAPI.h:109-110; FFTSVDpbeAPI.c:504-565) */
void writeMODPDB(const char* filename, PDBentry* PDBentries, unsigned int numPDBentries);
void writeFRC(const char* filename, PDBentry* PDBentries, unsigned int numPDBentries);

/* Molecule handling (reference FFTSVDpbeAPI.h:114-116; FFTSVDpbeAPI.c:569-672,
 * 1034-1049).  generateFlatSRF meshes the molecule with this library's SES
 * mesher (the reference runs its meshmaker/MSMS) and writes the .srf and
 * <srf>.d0.vert/.face; generateCurvedSRF reports that curved panels are not
 * supported. */
void assignRadiiCharges(PDBentry* PDBentries, unsigned int numPDBentries, SIZentry* SIZentries,
                        unsigned int numSIZentries, CRGentry* CRGentries, unsigned int numCRGentries);
void generateFlatSRF(const char* pdbfilename, const char* sizfilename, const char* srffilename);
void generateCurvedSRF(const char* pdbfilename, const char* sizfilename, const char* srffilename);

/* Surfaces: reads an .srf file and its .vert/.face surfaces
 * (reference FFTSVDpbeAPI.h:118-121; FFTSVDpbeAPI.c:776-1032).  Flat
 * panels only ('f' surfaces); curved surfaces are rejected with error(). */
void readSRF(const char* filename, Panel*** saltpanels, unsigned int** numsaltpanels, unsigned int* numsalts,
             Panel*** dielectricpanels, unsigned int** numdielectricpanels, unsigned int* numdielectrics,
             unsigned int** dielectricparent, Panel*** dielectriccavitypanels,
             unsigned int** numdielectriccavitypanels, unsigned int* numdielectriccavities,
             unsigned int** dielectriccavityparent, Panel*** saltcavitypanels, unsigned int** numsaltcavitypanels,
             unsigned int* numsaltcavities, unsigned int** saltcavityparent, unsigned int* numtotalpanels);

/* Operator (reference FFTSVDpbeAPI.h:125-131; FFTSVDpbeAPI.c:1053-1739).
 * Stern-layer formulation when ionexclusionradius > 0 with salt (salt
 * surfaces enclosing the dielectrics, dielectric cavities, salt cavities
 * inside those), otherwise the salt-without-Stern and no-salt
 * formulations. */
void generateSurfaceOperator(SurfaceOperator* pbesurfaceoperator, PDBentry* PDBentries, unsigned int numPDBentries,
                             Panel** saltpanels, unsigned int* numsaltpanels, unsigned int numsalts,
                             Panel** dielectricpanels, unsigned int* numdielectricpanels, unsigned int numdielectrics,
                             unsigned int* dielectricparent, Panel** dielectriccavitypanels,
                             unsigned int* numdielectriccavitypanels, unsigned int numdielectriccavities,
                             unsigned int* dielectriccavityparent, Panel** saltcavitypanels,
                             unsigned int* numsaltcavitypanels, unsigned int numsaltcavities,
                             unsigned int* saltcavityparent, unsigned int numtotalpanels);

/* reference FFTSVDpbeAPI.h:147; FFTSVDpbeAPI.c:1853-1861 */
void generateSurfaceOperatorPreconditioner(Preconditioner* preconditioner, SurfaceOperator pbesurfaceoperator,
                                           unsigned int numtotalpanels);
/* reference FFTSVDpbeAPI.h:150; FFTSVDpbeAPI.c:1871-1876 */
void generateRHS(Vector* rhs, SurfaceOperator pbesurfaceoperator, unsigned int numtotalpanels);
/* reference FFTSVDpbeAPI.h:153; FFTSVDpbeAPI.c:1884-1888 (GMRES to errortolerance) */
void solveSurfaceOperator(Vector* solution, SurfaceOperator pbesurfaceoperator, Preconditioner preconditioner,
                          Vector rhs, unsigned int numtotalpanels);
/* reference FFTSVDpbeAPI.h:156; FFTSVDpbeAPI.c:1896-1922 */
void collectPotentials(PDBentry* PDBcollectentries, unsigned int numPDBcollectentries, PDBentry* PDBinputentries,
                       unsigned int numPDBinputentries, SurfaceOperator pbesurfaceoperator, Vector solution);

/* Qualocation (second-kind, apparent-surface-charge) formulation, Laplace
 * only (reference FFTSVDpbeAPI.h:133-158; FFTSVDpbeAPI.c:1741-1869,1878-1882,
 * 1890-1894,1924-1950).  One tree over the dielectric and cavity panels with
 * double-layer-only bases; one unknown per panel.  error() when ionicstrength
 * is not 0.  With accelerateM3 the M1M3 panels -> charges tree carries the
 * right-hand side and the potentials (QualocationOperator.c:33-70);
 * otherwise they are direct sums (the reference's default). */
void generateQualocationOperator(QualocationOperator* qualocationoperator, PDBentry* PDBentries,
                                 unsigned int numPDBentries, Panel** dielectricpanels,
                                 unsigned int* numdielectricpanels, unsigned int numdielectrics,
                                 Panel** dielectriccavitypanels, unsigned int* numdielectriccavitypanels,
                                 unsigned int numdielectriccavities, unsigned int* dielectriccavityparent,
                                 unsigned int numtotalpanels);
/* The same operator; in the reference the tree is filled only by
 * generateQualocationOperator (FFTSVDpbeAPI.c:1781-1839), here both build it. */
void generateQualocationOperatorMinimal(QualocationOperator* qualocationoperator, PDBentry* PDBentries,
                                        unsigned int numPDBentries, Panel** dielectricpanels,
                                        unsigned int* numdielectricpanels, unsigned int numdielectrics,
                                        Panel** dielectriccavitypanels, unsigned int* numdielectriccavitypanels,
                                        unsigned int numdielectriccavities, unsigned int* dielectriccavityparent,
                                        unsigned int numtotalpanels);
/* reference FFTSVDpbeAPI.c:1863-1869: diagonal preconditioner */
void generateQualocationOperatorPreconditioner(Preconditioner* preconditioner,
                                               QualocationOperator qualocationoperator, unsigned int numtotalpanels);
/* reference FFTSVDpbeAPI.c:1878-1882 (Charge_makerhs_ecf_qual) */
void generateQualocationRHS(Vector* rhs, QualocationOperator qualocationoperator, unsigned int numtotalpanels);
/* reference FFTSVDpbeAPI.c:1890-1894 (GMRES_QualocationOperator to errortolerance;
 * the iteration cap is MAXITERATIONS = 1000) */
void solveQualocationOperator(Vector* solution, QualocationOperator qualocationoperator,
                              Preconditioner preconditioner, Vector rhs, unsigned int numtotalpanels);
/* reference FFTSVDpbeAPI.c:1924-1950 */
void collectQualocationPotentials(PDBentry* PDBcollectentries, unsigned int numPDBcollectentries,
                                  PDBentry* PDBinputentries, unsigned int numPDBinputentries,
                                  QualocationOperator qualocationoperator, Vector solution);
/* GAMESS entry (reference FFTSVDpbeAPI.h:159-163; FFTSVDpbeAPI.c:1974-2026):
 * panels as 9 doubles each (v1, v2, v3, used in that order), the qualocation
 * operator with 27 panels per leaf, 3 grid points, SVD 1e-5, GMRES to 1e-5;
 * sol receives one value per panel.  A zero right-hand side gives a zero
 * solution (the reference divides by its zero norm). */
void solveBEMproblem(unsigned int* numpanels, real* paneldata, real* innerdielectric, real* outerdielectric,
                     real* rhs, real* sol);

#ifdef __cplusplus
}
#endif
#endif
