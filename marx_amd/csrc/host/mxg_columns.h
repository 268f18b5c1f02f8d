/* mxg_columns.h -- the output columns of marx_write_photons, in the order of
 * the reference's table (marx/libsrc/marxio.c:259-289), shared by the host
 * writer (output.c) and the device packer (output.hip). */
#ifndef MXG_COLUMNS_H
#define MXG_COLUMNS_H

enum
{
   MXG_COL_B_ENERGY, MXG_COL_ENERGY, MXG_COL_TIME, MXG_COL_TAG,
   MXG_COL_XPOS, MXG_COL_YPOS, MXG_COL_ZPOS, MXG_COL_COSX, MXG_COL_COSY, MXG_COL_COSZ,
   MXG_COL_PHA, MXG_COL_CCDID, MXG_COL_CHIPX, MXG_COL_CHIPY, MXG_COL_U, MXG_COL_V,
   MXG_COL_MIRROR, MXG_COL_HRCREGION, MXG_COL_ORDER, MXG_COL_FINE, MXG_COL_COARSE1,
   MXG_COL_COARSE2, MXG_COL_COARSE3, MXG_COL_RA, MXG_COL_DEC, MXG_COL_ROLL,
   MXG_COL_DET_DY, MXG_COL_DET_DZ, MXG_COL_DET_THETA,
   MXG_NUM_COLS
};

/* bytes per value: 'E' and 'J' 4, 'I' 2, 'A' 1 */
#define MXG_COL_WIDTH(c) \
   (((c) == MXG_COL_PHA || (c) == MXG_COL_MIRROR) ? 2 \
    : ((c) == MXG_COL_CCDID || (c) == MXG_COL_HRCREGION || ((c) >= MXG_COL_ORDER && (c) <= MXG_COL_COARSE3)) ? 1 \
    : 4)

#endif
