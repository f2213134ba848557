/*
 * xmcml_io.h -- xmcml's file formats and its backup/resume loop, on the host.
 *
 * The reference launcher reads an XML <Input> description and a binary
 * .surface file, writes a sectioned binary .mcml.out result, and, every
 * backup portion, a .mcml.bk file holding the tallies so far plus the random
 * generator state from which a later run resumes.  These entry points read and
 * write the same byte layouts, so files made by the reference load here and
 * files written here load in the reference's readers and its visualisers:
 *
 *   ParseInputFile    (base/xmcml/xmcmlLauncher/parser.cpp:18-486)
 *       -> xmc_read_input_file()
 *   ParseSurfaceFile  (parser.cpp:555-587)
 *       -> xmc_read_surface_file()
 *   WriteOutputToFile (base/xmcml/xmcmlLauncher/writer.cpp:431-471)
 *       -> xmc_write_output_file()
 *   WriteBackupToFile (writer.cpp:473-522)
 *       -> xmc_write_backup_file()
 *   ParseBackupFile   (parser.cpp:589-611, with reader.cpp:27-392)
 *       -> xmc_read_backup_file()
 *   LaunchOMPWithBackups + DoBackup (base/xmcml/xmcmlLauncher/launcher_mpi.cpp:311-380)
 *       -> xmc_launch_with_backups()
 *   WriteWeightIntegralTableToFile / ParseWeightIntegralFile
 *                     (base/xmcml/xmcmlLauncher/weight_integral.cpp:105-247)
 *       -> xmc_write_weight_table_file() / xmc_read_weight_table_file()
 *
 * Differences from the reference, all deliberate:
 *   - Every function returns an int status (XMC_OK or a negative code, the
 *     message in xmc_last_error()).  The reference returns bool, or exits
 *     the process on a malformed file (parser.cpp:296-300).  Inputs that
 *     would make the reference read uninitialised memory or write past an
 *     array (a <Layers> element before <NumberOfLayers>, more <Id> than
 *     NumberOfSurfaces, fewer detector elements than declared, an element
 *     with no text where a number is expected) are rejected with XMC_ERR_ARG.
 *   - The random-generator section (0x09) of a backup holds one record with
 *     numThreadsPerProcess = numProcesses = 1, value = the stream seed and
 *     offset = the next photon index: the engine's streams are per photon,
 *     so (seed, next photon) is the whole state (xmcml.h, xmc_rng).
 *   - A NULL absorption or scattering array of the output is written as
 *     zeros; a NULL trajectory array as trajectorySize zeros, and skipped
 *     when reading.
 */
#ifndef XMCML_AMD_XMCML_IO_H
#define XMCML_AMD_XMCML_IO_H

#include "xmcml.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Section identifiers of the binary files, sections.h:4-19. */
#define XMC_SECTION_NUMBER_OF_PHOTONS 0x01
#define XMC_SECTION_AREA 0x02
#define XMC_SECTION_CUBE_DETECTORS 0x03
#define XMC_SECTION_SPECULAR_REFLECTANCE 0x04
#define XMC_SECTION_ABSORPTION 0x05
#define XMC_SECTION_DETECTOR_WEIGHTS 0x06
#define XMC_SECTION_DETECTOR_TRAJECTORIES 0x07
#define XMC_SECTION_DETECTOR_TIME_SCALE 0x08
#define XMC_SECTION_RANDOM_GENERATOR 0x09
#define XMC_SECTION_RING_DETECTORS 0x0A
#define XMC_SECTION_DETECTOR_RANGES 0x0B
#define XMC_SECTION_SCATTERING 0x0C
#define XMC_SECTION_SURFACES 0x0100
#define XMC_SECTION_WEIGHT_INTEGRAL 0x0200

/* The <Biasing> block of an input file (InputInfo.useBiasing..targetPoint,
 * mcml_kernel_types.h:74-80; defaults parser.cpp:10-12,32-41).  The same
 * values are set in the document's xmc_input, so a launch of it runs
 * attractive-vector biasing when the file asks for it. */
typedef struct xmc_biasing {
    int useBiasing;
    int weightIntegralPrecision; /* default 128 */
    int weightTablePrecision;    /* default 262144 */
    double attractiveFactor;     /* default 1.0 */
    xmc_double3 targetPoint;     /* read with atoi, as parser.cpp:233-241 does */
} xmc_biasing;

/* An input read from files: owns the xmc_input and every array it points to. */
typedef struct xmc_document xmc_document;

/* ParseInputFile: the input file's values with the reference's defaults
 * (timeScaleSize 1, time range 0..0, start (0,0,0) along +z, no detectors,
 * no target-range layers, terminationMode KILL).  Cube and ring detectors
 * are ordered by their number of filter layers, ties in file order
 * (parser.cpp:498-511).  The document has no surfaces until
 * xmc_read_surface_file(). */
int xmc_read_input_file(const char* fileName, xmc_document** doc);
/* ParseSurfaceFile: section 0x0100, then NumberOfSurfaces surfaces of
 * (int nV, nV x 3 doubles, int nT, nT x 3 ints); the count must equal the
 * input file's NumberOfSurfaces. */
int xmc_read_surface_file(xmc_document* doc, const char* fileName);
/* The document's input (valid until xmc_document_free); the caller may edit
 * its scalar fields, e.g. numberOfPhotons when resuming from a backup. */
xmc_input* xmc_document_input(xmc_document* doc);
int xmc_document_biasing(const xmc_document* doc, xmc_biasing* biasing);
void xmc_document_free(xmc_document* doc);

/* The .surface layout that xmc_read_surface_file reads (the reference ships
 * no writer; its surface files come from external tools). */
int xmc_write_surface_file(const xmc_surface* surfaces, int numberOfSurfaces, const char* fileName);

/* WriteOutputToFile: sections 0x01 (input->numberOfPhotons), 0x02, 0x03,
 * 0x0A, 0x04, 0x05, 0x0C, 0x06, 0x07, 0x08, 0x0B in that order. */
int xmc_write_output_file(const xmc_input* input, const xmc_output* output, const char* fileName);

/* WriteBackupToFile: the output file's sections followed by 0x09.
 * fileName NULL = the reference's name "xmcml_<count>.mcml.bk" in the
 * working directory, <count> = input->numberOfPhotons as writer.cpp:8-29
 * formats it (e.g. 1.50M, 20.00T for twenty thousand). */
int xmc_write_backup_file(const xmc_input* input, const xmc_output* output, const xmc_rng* rng,
                          const char* fileName);
/* The reference's backup file name for a photon count (buffer of >= 32 bytes). */
int xmc_backup_file_name(uint64_t numberOfPhotons, char* buffer, int bufferSize);

/* ParseBackupFile + ReadOutputFormBackupFile: output->numberOfPhotons,
 * absorption, scattering, detector weights, photon counts, trajectories,
 * time bins and target weights are replaced by the file's (the output must
 * have been made by xmc_init_output for the same input: grid size, detector
 * count and time-bin edges are checked); specularReflectance is not read
 * (reader.cpp:375-378 skips it).  rng receives (seed, next photon). */
int xmc_read_backup_file(const char* fileName, xmc_output* output, xmc_rng* rng);

/* The biasing's weight tables (the launcher's -t file): section 0x0200 with
 * the attractive factor, the integral precision, the number of tables and per
 * table (anisotropy, number of elements, elements).  The writer takes
 * input->weightTable (fileName NULL = the reference's
 * "xmcml_weight_integral_table.bin" in the working directory); the reader
 * checks the file against the document's input as ParseWeightIntegralFile
 * does (same attractive factor and precisions, a table within 1e-6 of every
 * distinct anisotropy of layers 1..) and makes the tables the input's
 * weightTable. */
int xmc_write_weight_table_file(const xmc_input* input, const char* fileName);
int xmc_read_weight_table_file(xmc_document* doc, const char* fileName);

/* LaunchOMPWithBackups + DoBackup for one process: runs input->numberOfPhotons
 * photons from rng->nextPhoton in portions of portionSize on one scene; after
 * every portion but the last it adds the tallies into output, advances
 * output->numberOfPhotons and rng->nextPhoton and writes a backup
 * (backupFileName, or the reference's name when NULL).  After the last portion
 * output->numberOfPhotons has grown by input->numberOfPhotons.  Resuming =
 * xmc_read_backup_file, then input->numberOfPhotons -= output->numberOfPhotons
 * (main.cpp:119-127), then this call again: photon p always draws stream
 * block p, so a resumed run adds the same photons as an uninterrupted one. */
int xmc_launch_with_backups(const xmc_input* input, xmc_output* output, xmc_rng* rng,
                            const xmc_options* options, uint64_t portionSize, const char* backupFileName);

#ifdef __cplusplus
}
#endif

#endif /* XMCML_AMD_XMCML_IO_H */
