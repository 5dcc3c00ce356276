/* prints the offsets and sizes of the drop-in's struct mirrors, one per line */
#include <stddef.h>
#include <stdio.h>

#include "../../include/gnx_refapi.h"
#include "layout_fields.h"

int main(void)
{
#define OFF(T, F) printf("%d\n", (int)offsetof(struct T, F));
#define SZ(T) printf("%d\n", (int)sizeof(struct T));
    GNX_LAYOUT_FIELDS(OFF)
    GNX_LAYOUT_TYPES(SZ)
    return 0;
}
