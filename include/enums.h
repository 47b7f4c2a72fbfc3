/* enums.h -- code rates, identical to the reference's include/enums.h:4-12. */
#ifndef ENUMS_H
#define ENUMS_H

enum coderate
{
    HALFRATE = 0,
    TWOTHIRDSA = 1,
    TWOTHIRDSB = 2,
    THREEQUARTERSA = 3,
    THREEQUARTERSB = 4,
    FIVESIXTHS = 5
};

#endif
