* Hand-made MPS file exercising the reader; its standard form is derived by
* hand in tests/test_mps.py.
NAME          EXAMPLE
ROWS
 N  COST
 L  LIM1
 G  LIM2
 E  MYEQN
 G  R4
 L  R5
 E  R6
 N  FREE2
COLUMNS
    X1        COST         1.0   LIM1         1.0
    X1        LIM2         1.0
    MARKER                 'MARKER'                 'INTORG'
    X2        COST         2.0   LIM1         1.0
    X2        MYEQN       -1.0
    MARKER                 'MARKER'                 'INTEND'
    X3        COST        -1.0   MYEQN        1.0
    X3        R4           2.0   FREE2        5.0
    X4        COST         0.5   R5           3.0
    X4        R5           1.0   R6           1.0
    X4        LIM2         2.0
    X5        COST         1.0   LIM2        -1.0
    X6        R5           1.0
    X7        MYEQN        1.0
RHS
    RHS       COST        -7.5
    RHS       LIM1         4.0   LIM2         1.0
    RHS       MYEQN        7.0   R4           2.0
    RHS       R5           9.0   R6           3.0
RANGES
    RNG       LIM1         2.5   R4           3.0
    RNG       R6          -2.0
BOUNDS
 UP BND       X1           4.0
 LO BND       X2          -1.0
 UP BND       X2           1.0
 FR BND       X3
 MI BND       X4
 UP BND       X4          -2.0
 UP BND       X5          -3.0
 BV BND       X6
 FX BND       X7           2.0
ENDATA
