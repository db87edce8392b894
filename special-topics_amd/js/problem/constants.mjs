// SI constants used by the FP64 problem setup.
//
// The literal values (and the expression for mu0) are the ones js_fdtd uses
// (js_fdtd/src/constants.js:3-10); they must be identical, bit for bit, or every
// coefficient array downstream changes in its last bits.

export const EPS0 = 8.854187817e-12;        // F/m
export const MU0 = 4 * Math.PI * 1e-7;      // H/m, evaluated as (4*pi)*1e-7
export const C0 = 2.99792458e8;             // m/s (literal, not 1/sqrt(mu0*eps0))
export const ETA0 = 376.73031346177066;     // ohm
export const Q_E = 1.602e-19;               // C
export const K_B = 1.38066e-23;             // J/K
